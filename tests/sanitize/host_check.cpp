// Host-side checks of libmxb_diffusion built with AddressSanitizer and
// UndefinedBehaviorSanitizer on its host code (tests/sanitize/build.py).
//
//   host_check cpu   argument validation and the no-device error paths
//   host_check gpu   every host-side driver on small volumes: in-memory,
//                    z-chunked (threaded staged copies), file-backed (mmap'd
//                    state files), the in-memory -> chunked fallback, batched,
//                    device-resident and no-allocation forms, kernel timing,
//                    the field and EDT entries; results cross-checked bit for
//                    bit where the forms promise it
//
// Exit status 0 when every check passes; a sanitizer report aborts the run.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include <unistd.h>

#include "maxibone_diffusion.h"
#include "maxibone_edt.h"
#include "maxibone_fields.h"
#include "maxibone_probe.h"

static int g_failures = 0;

#define EXPECT(cond)                                                                              \
    do {                                                                                          \
        if (!(cond)) {                                                                            \
            std::fprintf(stderr, "FAIL %s:%d  %s  (last error: %s)\n", __FILE__, __LINE__, #cond, \
                         mxb_last_error());                                                       \
            ++g_failures;                                                                         \
        }                                                                                         \
    } while (0)

static std::vector<float> gauss_taps(double sigma) {
    const int r = (int)std::lround(4.0 * sigma);
    std::vector<double> d(2 * r + 1);
    double s = 0;
    for (int i = -r; i <= r; ++i) s += d[i + r] = std::exp(-0.5 / (sigma * sigma) * (double)(i * i));
    std::vector<float> k(d.size());
    for (size_t i = 0; i < d.size(); ++i) k[i] = (float)(d[i] / s);
    return k;
}

static std::vector<uint8_t> spheres(int64_t nz, int64_t ny, int64_t nx, unsigned seed) {
    std::vector<uint8_t> m((size_t)(nz * ny * nx), 0);
    std::mt19937 g(seed);
    for (int s = 0; s < 6; ++s) {
        const double cz = g() % nz, cy = g() % ny, cx = g() % nx, r = 2 + g() % 6;
        for (int64_t z = 0; z < nz; ++z)
            for (int64_t y = 0; y < ny; ++y)
                for (int64_t x = 0; x < nx; ++x)
                    if ((z - cz) * (z - cz) + (y - cy) * (y - cy) + (x - cx) * (x - cx) <= r * r)
                        m[(size_t)((z * ny + y) * nx + x)] = 1;
    }
    return m;
}

template <class T>
static bool same(const std::vector<T> &a, const std::vector<T> &b) {
    return a.size() == b.size() && std::memcmp(a.data(), b.data(), a.size() * sizeof(T)) == 0;
}

static std::string temp_path(const char *tag) {
    const char *dir = std::getenv("TMPDIR");
    std::string p = std::string(dir ? dir : "/tmp") + "/mxb_host_check_" + tag + "_XXXXXX";
    std::vector<char> buf(p.begin(), p.end());
    buf.push_back('\0');
    const int fd = mkstemp(buf.data());
    if (fd >= 0) close(fd);
    return std::string(buf.data());
}

static int cpu_checks() {
    const std::vector<float> k = gauss_taps(1.0);
    const int64_t K = (int64_t)k.size();
    std::vector<uint8_t> m(4 * 5 * 6, 0);
    std::vector<uint16_t> q(m.size());
    std::vector<float> u(m.size());

    EXPECT(mxb_version() != nullptr && std::strncmp(mxb_version(), "maxibone_amd", 12) == 0);
    EXPECT(mxb_state_bytes(MXB_FP32, 4, 5, 6) == 512 + 2048 * 4 + 80);  // + the split form's low-plane gap
    EXPECT(mxb_state_bytes(1024, 4, 5, 6) == -1);
    EXPECT(mxb_state_bytes(MXB_FP32, 4, -5, 6) == -1);
    EXPECT(std::strlen(mxb_last_error()) > 0);

    const std::vector<float> even(4, 0.25f);
    EXPECT(mxb_diffusion_u8_u16(m.data(), 4, 5, 6, even.data(), 4, 1, q.data(), 0) == MXB_EINVAL);
    EXPECT(mxb_diffusion_u8_u16(nullptr, 4, 5, 6, k.data(), K, 1, q.data(), 0) == MXB_EINVAL);
    EXPECT(mxb_diffusion_u8_u16(m.data(), 4, 5, 6, k.data(), K, -1, q.data(), 0) == MXB_EINVAL);
    EXPECT(mxb_diffusion_ex(m.data(), 4, 5, 6, k.data(), K, 1, 99, q.data(), nullptr, 0) == MXB_EINVAL);
    EXPECT(mxb_diffusion_ex(m.data(), 4, 5, 6, k.data(), K, 1, MXB_FP32, nullptr, nullptr, 0) == MXB_EINVAL);
    EXPECT(mxb_diffusion_chunked(m.data(), 4, 5, 6, k.data(), K, 1, 0, 1, MXB_FP32, q.data(), nullptr, 0) ==
           MXB_EINVAL);
    EXPECT(mxb_diffusion_chunked(m.data(), 4, 5, 6, k.data(), K, 1, 2, 0, MXB_FP32, q.data(), nullptr, 0) ==
           MXB_EINVAL);
    EXPECT(mxb_diffusion_on_disk_ex("/nonexistent/in.u8", k.data(), K, "/nonexistent/out.u16", 4, 5, 6, 2, 5, 6, 1,
                                    MXB_FP32, 7, 0) == MXB_EINVAL);  // unknown flag
    EXPECT(mxb_diffusion_on_disk("/nonexistent/in.u8", k.data(), K, "/nonexistent/out.u16", 4, 5, 6, 2, 5, 7, 1,
                                 0) == MXB_EINVAL);  // chunk x differs from the volume's
    EXPECT(mxb_device_diffusion_noalloc(nullptr, 512, 512, 512, k.data(), K, 1, MXB_FP32, nullptr, nullptr,
                                        nullptr) == MXB_EINVAL);
    EXPECT(mxb_kernel_times(99, nullptr, nullptr) == MXB_EINVAL);
    float ms = 0;
    EXPECT(mxb_probe_hbm(MXB_PROBE_COPY, nullptr, nullptr, 100, 1, nullptr, &ms) == MXB_EINVAL);
    EXPECT(mxb_batch_create(nullptr, 0, 1, MXB_FP32, nullptr, nullptr) == MXB_EINVAL);

    // with no device the valid calls fail with MXB_EDEVICE, not a crash
    if (mxb_device_count() == 0) {
        EXPECT(mxb_diffusion_u8_u16(m.data(), 4, 5, 6, k.data(), K, 1, q.data(), 0) == MXB_EDEVICE);
        EXPECT(mxb_diffusion_chunked(m.data(), 4, 5, 6, k.data(), K, 2, 2, 1, MXB_BF16, q.data(), u.data(), 0) ==
               MXB_EDEVICE);
    }
    return g_failures;
}

static int gpu_checks() {
    if (mxb_device_count() < 1) {
        std::fprintf(stderr, "no device\n");
        return 1;
    }
    // x rows of 70 voxels: neither 128-byte nor 16-byte aligned rows
    const int64_t nz = 37, ny = 45, nx = 70, n = nz * ny * nx, reps = 5;
    const std::vector<float> k = gauss_taps(1.5);
    const int64_t K = (int64_t)k.size();
    const std::vector<uint8_t> m = spheres(nz, ny, nx, 7);
    std::vector<uint16_t> q_ref(n), q(n);
    std::vector<float> u_ref(n), u(n);

    std::printf("in-memory fp32\n");
    EXPECT(mxb_diffusion_ex(m.data(), nz, ny, nx, k.data(), K, reps, MXB_FP32, q_ref.data(), u_ref.data(), 0) ==
           MXB_OK);
    EXPECT(mxb_diffusion_u8_u16(m.data(), nz, ny, nx, k.data(), K, reps, q.data(), 0) == MXB_OK && same(q, q_ref));

    std::printf("chunked fp32 (z = 8, temporal 2)\n");
    EXPECT(mxb_diffusion_chunked(m.data(), nz, ny, nx, k.data(), K, reps, 8, 2, MXB_FP32, q.data(), u.data(), 0) ==
           MXB_OK);
    EXPECT(same(q, q_ref) && same(u, u_ref));

    std::printf("chunked bf16 vs in-memory bf16\n");
    std::vector<uint16_t> qb(n);
    EXPECT(mxb_diffusion_ex(m.data(), nz, ny, nx, k.data(), K, reps, MXB_BF16, qb.data(), nullptr, 0) == MXB_OK);
    EXPECT(mxb_diffusion_chunked(m.data(), nz, ny, nx, k.data(), K, reps, 11, 1, MXB_BF16, q.data(), nullptr, 0) ==
           MXB_OK);
    EXPECT(same(q, qb));

    std::printf("in-memory -> chunked fallback (MXB_DEVICE_BUDGET)\n");
    setenv("MXB_DEVICE_BUDGET", "200000", 1);
    EXPECT(mxb_diffusion_u8_u16(m.data(), nz, ny, nx, k.data(), K, reps, q.data(), 0) == MXB_OK);
    unsetenv("MXB_DEVICE_BUDGET");
    EXPECT(same(q, q_ref));

    std::printf("file-backed (z chunks of 10)\n");
    const std::string fin = temp_path("in"), fout = temp_path("out");
    if (FILE *f = std::fopen(fin.c_str(), "wb")) {
        EXPECT(std::fwrite(m.data(), 1, m.size(), f) == m.size());
        std::fclose(f);
    }
    EXPECT(mxb_diffusion_on_disk(fin.c_str(), k.data(), K, fout.c_str(), nz, ny, nx, 10, ny, nx, reps, 0) ==
           MXB_OK);
    std::fill(q.begin(), q.end(), 0);
    if (FILE *f = std::fopen(fout.c_str(), "rb")) {
        EXPECT(std::fread(q.data(), 2, (size_t)n, f) == (size_t)n);
        std::fclose(f);
    }
    EXPECT(same(q, q_ref));
    EXPECT(mxb_diffusion_on_disk_ex(fin.c_str(), k.data(), K, fout.c_str(), nz, ny, nx, 10, ny, nx, reps, MXB_FP32,
                                    MXB_ON_DISK_REFERENCE_MASK, 0) == MXB_OK);
    std::remove(fin.c_str());
    std::remove(fout.c_str());

    std::printf("batched (three volumes) vs single-volume calls\n");
    const int64_t shapes[3][3] = {{20, 33, 41}, {9, 64, 64}, {30, 17, 90}};
    std::vector<std::vector<uint8_t>> bm;
    std::vector<std::vector<uint16_t>> bq(3), bref(3);
    std::vector<mxb_volume_desc> desc(3);
    for (int i = 0; i < 3; ++i) {
        const int64_t *s = shapes[i];
        bm.push_back(spheres(s[0], s[1], s[2], 11 + i));
        bq[i].resize((size_t)(s[0] * s[1] * s[2]));
        bref[i].resize(bq[i].size());
        desc[i] = mxb_volume_desc{bm[i].data(), s[0], s[1], s[2], k.data(), K, bq[i].data(), nullptr};
        EXPECT(mxb_diffusion_ex(bm[i].data(), s[0], s[1], s[2], k.data(), K, 4, MXB_FP32, bref[i].data(), nullptr,
                                0) == MXB_OK);
    }
    EXPECT(mxb_diffusion_batched(desc.data(), 3, 4, MXB_FP32, 0) == MXB_OK);
    for (int i = 0; i < 3; ++i) EXPECT(same(bq[i], bref[i]));

    std::printf("device-resident and no-allocation forms, kernel timing\n");
    void *d_mask = nullptr, *d_s0 = nullptr, *d_s1 = nullptr, *d_q = nullptr, *d_q2 = nullptr;
    const int64_t sb = mxb_state_bytes(MXB_FP32, nz, ny, nx);
    EXPECT(sb > 0);
    EXPECT(hipMalloc(&d_mask, (size_t)n) == hipSuccess && hipMalloc(&d_s0, (size_t)sb) == hipSuccess &&
           hipMalloc(&d_s1, (size_t)sb) == hipSuccess && hipMalloc(&d_q, 2 * (size_t)n) == hipSuccess &&
           hipMalloc(&d_q2, 2 * (size_t)n) == hipSuccess);
    EXPECT(hipMemcpy(d_mask, m.data(), (size_t)n, hipMemcpyHostToDevice) == hipSuccess);
    EXPECT(mxb_set_kernel_timing(1) == MXB_OK);
    EXPECT(mxb_device_diffusion(static_cast<const uint8_t *>(d_mask), nz, ny, nx, k.data(), K, reps, MXB_FP32, d_s0,
                                d_s1, static_cast<uint16_t *>(d_q), nullptr, nullptr) == MXB_OK);
    EXPECT(hipDeviceSynchronize() == hipSuccess);
    double total = 0;
    int64_t launches = 0;  // read before switching timing off, which resets the totals
    EXPECT(mxb_kernel_times(MXB_K_ZPASS, &total, &launches) == MXB_OK && launches >= reps && total > 0);
    EXPECT(mxb_set_kernel_timing(0) == MXB_OK);
    EXPECT(mxb_device_diffusion_noalloc(static_cast<const uint8_t *>(d_mask), nz, ny, nx, k.data(), K, reps,
                                        MXB_FP32, static_cast<uint16_t *>(d_q2), nullptr, nullptr) == MXB_OK);
    EXPECT(hipDeviceSynchronize() == hipSuccess);
    EXPECT(hipMemcpy(q.data(), d_q, 2 * (size_t)n, hipMemcpyDeviceToHost) == hipSuccess && same(q, q_ref));
    EXPECT(hipMemcpy(q.data(), d_q2, 2 * (size_t)n, hipMemcpyDeviceToHost) == hipSuccess && same(q, q_ref));
    for (void *p : {d_mask, d_s0, d_s1, d_q, d_q2}) hipFree(p);

    std::printf("field and EDT entries\n");
    std::vector<uint8_t> cyl((size_t)(ny * nx), 1);
    std::vector<uint16_t> edt(n), comb(n);
    EXPECT(mxb_edt_field(m.data(), nz, ny, nx, cyl.data(), edt.data(), 0) == MXB_OK);
    EXPECT(mxb_combine_fields(q_ref.data(), edt.data(), nz, ny, nx, cyl.data(), comb.data(), 0) == MXB_OK);
    const int64_t vb = 64, fb = 32;
    std::vector<uint64_t> bins((size_t)(vb * fb), 0);
    EXPECT(mxb_field_histogram(q_ref.data(), nz, ny, nx, comb.data(), nz, ny, nx, bins.data(), fb, vb, 1.0, 65535.0,
                               1.0, 65535.0, 0) == MXB_OK);
    return g_failures;
}

int main(int argc, char **argv) {
    const std::string mode = argc > 1 ? argv[1] : "cpu";
    const int f = mode == "gpu" ? gpu_checks() : cpu_checks();
    std::printf("%s checks: %d failure(s)\n", mode.c_str(), f);
    return f == 0 ? 0 : 1;
}
