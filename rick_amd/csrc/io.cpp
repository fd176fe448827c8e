// Data formats on either side of the imaging chain: the binMS measurement
// layout the reference reads (one raw array per quantity plus meta.txt) and
// the raw image files it writes.
//
//   rick_binms_read_meta   readMetaData          init.c:300-333
//   rick_binms_partition   metaData_calculation  init.c:335-372
//   rick_binms_read        readData              init.c:424-490
//   rick_binms_write       the writer of scripts/create_binMS.py:210-245
//   rick_write_image       fourier_transform.c:280-289
//   rick_write_cube        WRITE_DATA cubes, gridding.c:210-282 and
//                          fourier_transform.c:134-201 (single rank)
//
// Reads into device memory stream each file through two pinned staging
// buffers: while one chunk is copied to the GPU the next is read from disk.
#include <hip/hip_runtime.h>

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <algorithm>

#include "rick_internal.h"

using namespace rick;

namespace {

// File names of the reference's paramfile (paramfile.txt:11-17).
const char *const kFiles[6] = {"ucoord.bin", "vcoord.bin", "wcoord.bin", "weights.bin",
                               "visibilities_real.bin", "visibilities_img.bin"};
const char *const kMeta = "meta.txt";

std::string join(const char *dir, const char *name)
{
    std::string p = dir ? dir : "";
    if (!p.empty() && p.back() != '/') p += '/';
    return p + name;
}

constexpr size_t kChunk = 32ull << 20;  // bytes per staging buffer

struct Staging {
    void *buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipStream_t s = nullptr;
    ~Staging()
    {
        for (int i = 0; i < 2; ++i) {
            if (ev[i]) (void)hipEventDestroy(ev[i]);
            if (buf[i]) (void)hipHostFree(buf[i]);
        }
        if (s) (void)hipStreamDestroy(s);
    }
    int init(int device)
    {
        if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&s) != hipSuccess)
            return set_status(RICK_ERR_DEVICE, "cannot use device %d", device);
        for (int i = 0; i < 2; ++i) {
            if (hipHostMalloc(&buf[i], kChunk, hipHostMallocDefault) != hipSuccess)
                return set_status(RICK_ERR_NOMEM, "pinned staging allocation failed");
            if (hipEventCreateWithFlags(&ev[i], hipEventDisableTiming) != hipSuccess)
                return set_status(RICK_ERR_DEVICE, "event creation failed");
        }
        return RICK_OK;
    }
};

// Reads bytes [off, off+n) of path into dst: a host pointer, or a device
// pointer when st is non-null (streamed through the staging buffers).
int read_range(const std::string &path, size_t off, size_t n, void *dst, Staging *st)
{
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return set_status(RICK_ERR_INVALID, "cannot open %s: %s", path.c_str(), strerror(errno));
    if (std::fseek(f, (long)off, SEEK_SET) != 0) {
        std::fclose(f);
        return set_status(RICK_ERR_INVALID, "cannot seek %s to %zu", path.c_str(), off);
    }
    int rc = RICK_OK;
    if (!st) {
        if (n && std::fread(dst, 1, n, f) != n)
            rc = set_status(RICK_ERR_INVALID, "%s: short read (%zu bytes from %zu)", path.c_str(),
                            n, off);
    } else {
        size_t done = 0;
        int k = 0;
        while (done < n && rc == RICK_OK) {
            const size_t m = n - done < kChunk ? n - done : kChunk;
            // Buffer k was last handed to a copy two chunks ago: wait for it.
            if (hipEventSynchronize(st->ev[k]) != hipSuccess) {
                rc = set_status(RICK_ERR_DEVICE, "staging copy failed");
                break;
            }
            if (std::fread(st->buf[k], 1, m, f) != m) {
                rc = set_status(RICK_ERR_INVALID, "%s: short read at %zu", path.c_str(), off + done);
                break;
            }
            if (hipMemcpyAsync(static_cast<char *>(dst) + done, st->buf[k], m,
                               hipMemcpyHostToDevice, st->s) != hipSuccess ||
                hipEventRecord(st->ev[k], st->s) != hipSuccess) {
                rc = set_status(RICK_ERR_DEVICE, "host-to-device copy failed");
                break;
            }
            done += m;
            k ^= 1;
        }
        if (hipStreamSynchronize(st->s) != hipSuccess && rc == RICK_OK)
            rc = set_status(RICK_ERR_DEVICE, "host-to-device copy failed");
    }
    std::fclose(f);
    return rc;
}

int write_file(const std::string &path, const void *src, size_t n)
{
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return set_status(RICK_ERR_INVALID, "cannot create %s: %s", path.c_str(), strerror(errno));
    const bool ok = (n == 0 || std::fwrite(src, 1, n, f) == n);
    const bool closed = std::fclose(f) == 0;
    return ok && closed ? RICK_OK : set_status(RICK_ERR_INVALID, "cannot write %s", path.c_str());
}

}  // namespace

extern "C" {

int rick_binms_read_meta(const char *dir, rick_binms_meta *meta)
{
    clear_status();
    if (!meta) return set_status(RICK_ERR_INVALID, "null meta");
    const std::string path = join(dir, kMeta);
    FILE *f = std::fopen(path.c_str(), "r");
    if (!f) return set_status(RICK_ERR_INVALID, "cannot open %s", path.c_str());
    // The first twelve values, in the order and with the conversions of
    // init.c:309-320; further lines (the writer adds three) are ignored.
    rick_binms_meta m{};
    int got = 0;
    got += std::fscanf(f, "%u", &m.nmeasures) == 1;
    got += std::fscanf(f, "%lu", &m.nvis) == 1;
    got += std::fscanf(f, "%u", &m.freq_per_chan) == 1;
    got += std::fscanf(f, "%u", &m.polarisations) == 1;
    got += std::fscanf(f, "%u", &m.ntimes) == 1;
    got += std::fscanf(f, "%lf", &m.dt) == 1;
    got += std::fscanf(f, "%lf", &m.thours) == 1;
    got += std::fscanf(f, "%u", &m.baselines) == 1;
    got += std::fscanf(f, "%lf", &m.uvmin) == 1;
    got += std::fscanf(f, "%lf", &m.uvmax) == 1;
    got += std::fscanf(f, "%lf", &m.wmin) == 1;
    got += std::fscanf(f, "%lf", &m.wmax) == 1;
    std::fclose(f);
    if (got != 12) return set_status(RICK_ERR_INVALID, "%s: %d of 12 values parsed", path.c_str(), got);
    *meta = m;
    return RICK_OK;
}

int rick_binms_partition(const rick_binms_meta *meta, int rank, int size, unsigned int drop,
                         unsigned long *startrow, unsigned int *nrows)
{
    clear_status();
    if (!meta || !startrow || !nrows) return set_status(RICK_ERR_INVALID, "null argument");
    if (size < 1 || rank < 0 || rank >= size)
        return set_status(RICK_ERR_INVALID, "rank %d of %d", rank, size);
    if (meta->nmeasures < drop)
        return set_status(RICK_ERR_INVALID, "%u measurements, fewer than the %u dropped",
                          meta->nmeasures, drop);
    // init.c:337-359: the last `drop` rows are left out, the rest is split in
    // equal parts with the remainder on the last rank.
    const unsigned int total = meta->nmeasures - drop;
    unsigned int per = total / (unsigned int)size;
    const unsigned int rem = total % (unsigned int)size;
    *startrow = (unsigned long)rank * per;
    if (rank == size - 1) per += rem;
    *nrows = per;
    return RICK_OK;
}

int rick_binms_read(const char *dir, unsigned int freq_per_chan, unsigned int polarizations,
                    unsigned long startrow, unsigned int nrows, double *uu, double *vv, double *ww,
                    float *weight, float *vis_real, float *vis_img, int device)
{
    clear_status();
    if (freq_per_chan == 0 || polarizations == 0)
        return set_status(RICK_ERR_INVALID, "freq_per_chan and polarizations must be >= 1");
    if (nrows == 0) return RICK_OK;
    if (!uu || !vv || !ww || !weight || !vis_real || !vis_img)
        return set_status(RICK_ERR_INVALID, "null array");
    const size_t F = freq_per_chan, P = polarizations, r0 = startrow, n = nrows;
    // Byte offsets and sizes of init.c:429-488.
    void *dst[6] = {uu, vv, ww, weight, vis_real, vis_img};
    const size_t off[6] = {r0 * 8, r0 * 8, r0 * 8, r0 * P * 4, r0 * F * P * 4, r0 * F * P * 4};
    const size_t len[6] = {n * 8, n * 8, n * 8, n * P * 4, n * F * P * 4, n * F * P * 4};
    Staging st;
    Staging *sp = nullptr;
    int prev = 0, rc = RICK_OK;
    if (device >= 0) {
        (void)hipGetDevice(&prev);
        rc = st.init(device);
        sp = &st;
    }
    for (int i = 0; i < 6 && rc == RICK_OK; ++i)
        rc = read_range(join(dir, kFiles[i]), off[i], len[i], dst[i], sp);
    if (device >= 0) (void)hipSetDevice(prev);
    return rc;
}

int rick_binms_write(const char *dir, const rick_binms_meta *meta, const double *uu,
                     const double *vv, const double *ww, const float *weight,
                     const float *vis_real, const float *vis_img)
{
    clear_status();
    if (!meta || !uu || !vv || !ww || !weight || !vis_real || !vis_img)
        return set_status(RICK_ERR_INVALID, "null argument");
    const size_t n = meta->nmeasures, P = meta->polarisations, F = meta->freq_per_chan;
    if ((size_t)meta->nvis != n * F * P)
        return set_status(RICK_ERR_INVALID, "nvis %lu != nmeasures*freq*pol", meta->nvis);
    const void *src[6] = {uu, vv, ww, weight, vis_real, vis_img};
    const size_t len[6] = {n * 8, n * 8, n * 8, n * P * 4, n * F * P * 4, n * F * P * 4};
    for (int i = 0; i < 6; ++i) {
        const int r = write_file(join(dir, kFiles[i]), src[i], len[i]);
        if (r != RICK_OK) return r;
    }
    // meta.txt as create_binMS.py:225-245 writes it: the twelve values the
    // reader parses, then the weight array's shape (rows, polarizations, 1.0).
    const std::string path = join(dir, kMeta);
    FILE *f = std::fopen(path.c_str(), "w");
    if (!f) return set_status(RICK_ERR_INVALID, "cannot create %s", path.c_str());
    std::fprintf(f, "%u\n%lu\n%u\n%u\n%u\n%.17g\n%.17g\n%u\n%.17g\n%.17g\n%.17g\n%.17g\n%u\n%u\n1.0\n",
                 meta->nmeasures, meta->nvis, meta->freq_per_chan, meta->polarisations,
                 meta->ntimes, meta->dt, meta->thours, meta->baselines, meta->uvmin, meta->uvmax,
                 meta->wmin, meta->wmax, meta->nmeasures, meta->polarisations);
    return std::fclose(f) == 0 ? RICK_OK : set_status(RICK_ERR_INVALID, "cannot write %s", path.c_str());
}

int rick_write_image(const char *real_path, const char *imag_path, const double *image_real,
                     const double *image_imag, int xaxis, int yaxis, int rank)
{
    clear_status();
    if (!real_path || !imag_path || !image_real || !image_imag)
        return set_status(RICK_ERR_INVALID, "null argument");
    if (xaxis <= 0 || yaxis <= 0 || rank < 0) return set_status(RICK_ERR_INVALID, "bad shape");
    // fourier_transform.c:280-289: each file opened "wb", the slab written at
    // byte rank*xaxis*yaxis*8.
    const size_t n = (size_t)xaxis * (size_t)yaxis;
    const long at = (long)((size_t)rank * n * 8);
    const char *paths[2] = {real_path, imag_path};
    const double *src[2] = {image_real, image_imag};
    for (int i = 0; i < 2; ++i) {
        FILE *f = std::fopen(paths[i], "wb");
        if (!f) return set_status(RICK_ERR_INVALID, "cannot create %s", paths[i]);
        bool ok = std::fseek(f, at, SEEK_SET) == 0 && std::fwrite(src[i], sizeof(double), n, f) == n;
        ok = (std::fclose(f) == 0) && ok;
        if (!ok) return set_status(RICK_ERR_INVALID, "cannot write %s", paths[i]);
    }
    return RICK_OK;
}

int rick_write_cube(const char *real_path, const char *imag_path, const void *grid, int precision,
                    int num_w_planes, int grid_size_x, int grid_size_y, int device)
{
    clear_status();
    if (!real_path || !imag_path || !grid) return set_status(RICK_ERR_INVALID, "null argument");
    if (num_w_planes < 1 || grid_size_x < 1 || grid_size_y < 1)
        return set_status(RICK_ERR_INVALID, "bad shape");
    if (precision != RICK_FP64 && precision != RICK_FP32)
        return set_status(RICK_ERR_INVALID, "precision %d", precision);
    // One rank holds the whole grid, so the per-element seeks of the
    // reference (index iu + iv*Nx + iw*Nx*Ny) write the planes in order: the
    // files are the real and imaginary parts of the grid, as doubles.
    const size_t n = (size_t)num_w_planes * grid_size_x * grid_size_y;
    FILE *fr = std::fopen(real_path, "wb");
    FILE *fi = fr ? std::fopen(imag_path, "wb") : nullptr;
    if (!fr || !fi) {
        if (fr) std::fclose(fr);
        return set_status(RICK_ERR_INVALID, "cannot create %s / %s", real_path, imag_path);
    }
    int rc = RICK_OK;
    const size_t chunk = kChunk / sizeof(double);
    if (device < 0) {
        std::vector<double> re(std::min(n, chunk)), im(std::min(n, chunk));
        for (size_t at = 0; at < n && rc == RICK_OK; at += chunk) {
            const size_t m = std::min(chunk, n - at);
            for (size_t i = 0; i < m; ++i) {
                if (precision == RICK_FP32) {
                    const float *g = static_cast<const float *>(grid) + 2 * (at + i);
                    re[i] = g[0];
                    im[i] = g[1];
                } else {
                    const double *g = static_cast<const double *>(grid) + 2 * (at + i);
                    re[i] = g[0];
                    im[i] = g[1];
                }
            }
            if (std::fwrite(re.data(), sizeof(double), m, fr) != m ||
                std::fwrite(im.data(), sizeof(double), m, fi) != m)
                rc = set_status(RICK_ERR_INVALID, "cannot write the cube files");
        }
    } else {
        // De-interleave on the device a chunk at a time, copy both halves to
        // pinned host memory, write.
        int prev = 0;
        (void)hipGetDevice(&prev);
        Staging st;
        double *dre = nullptr;
        rc = st.init(device);
        if (rc == RICK_OK && hipMalloc(&dre, 2 * kChunk) != hipSuccess)
            rc = set_status(RICK_ERR_NOMEM, "device staging allocation failed");
        const size_t cb = precision == RICK_FP32 ? 8 : 16;
        for (size_t at = 0; at < n && rc == RICK_OK; at += chunk) {
            const size_t m = std::min(chunk, n - at);
            const void *src = static_cast<const char *>(grid) + at * cb;
            double *hre = static_cast<double *>(st.buf[0]), *him = static_cast<double *>(st.buf[1]);
            if (launch_split_complex(src, precision, m, dre, dre + chunk, st.s) != hipSuccess ||
                hipMemcpyAsync(hre, dre, m * 8, hipMemcpyDeviceToHost, st.s) != hipSuccess ||
                hipMemcpyAsync(him, dre + chunk, m * 8, hipMemcpyDeviceToHost, st.s) != hipSuccess ||
                hipStreamSynchronize(st.s) != hipSuccess) {
                rc = set_status(RICK_ERR_DEVICE, "device-to-host copy of the grid failed");
                break;
            }
            if (std::fwrite(hre, sizeof(double), m, fr) != m ||
                std::fwrite(him, sizeof(double), m, fi) != m)
                rc = set_status(RICK_ERR_INVALID, "cannot write the cube files");
        }
        if (dre) (void)hipFree(dre);
        (void)hipSetDevice(prev);
    }
    const int cr = std::fclose(fr), ci = std::fclose(fi);
    const bool closed = cr == 0 && ci == 0;
    if (rc == RICK_OK && !closed) rc = set_status(RICK_ERR_INVALID, "cannot write the cube files");
    return rc;
}

}  // extern "C"
