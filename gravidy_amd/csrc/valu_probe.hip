// valu_probe.hip -- issue cost of the VALU instruction kinds the pair
// kernels are made of, on the current GPU.  Each kernel runs 8 independent
// dependency chains per lane, 2048 workgroups of 256 threads (8 waves per
// SIMD on a 256-CU part), all in registers; the report is nanoseconds per
// chain step per SIMD, i.e. per wave64 instruction group of the listed
// kind.  Used to explain the mixed-precision pair's cost; not part of the
// library.
#include <hip/hip_runtime.h>

#include <cstdio>

namespace {

constexpr int ITERS = 4096;
constexpr int BLOCKS = 2048;

template <int OP>
__global__ __launch_bounds__(256) void probe(double *outd, float *outf, double a, float af,
                                             int iters) {
    double d[8];
    float f[8];
    for (int k = 0; k < 8; ++k) {
        d[k] = threadIdx.x * 1e-3 + k;
        f[k] = (float)d[k];
    }
    const double b = a * 0.5;
    const float bf = af * 0.5f;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if constexpr (OP == 0) d[k] = fma(d[k], a, b);
            if constexpr (OP == 1) d[k] = d[k] + a;
            if constexpr (OP == 2) f[k] = fmaf(f[k], af, bf);
            if constexpr (OP == 3) f[k] = f[k] + af;
            if constexpr (OP == 4) {
                f[k] = (float)d[k];
                d[k] = (double)f[k] * a;
            }
            if constexpr (OP == 5) d[k] = __builtin_amdgcn_rsq(d[k]) + a;
            if constexpr (OP == 6) f[k] = __builtin_amdgcn_rsqf(f[k]) + af;
            if constexpr (OP == 7) {
                d[k] = fma(d[k], a, b);
                f[k] = fmaf(f[k], af, bf);
            }
            if constexpr (OP == 8) {
                d[k] = d[k] + a;
                f[k] = fmaf(f[k], af, (float)d[k]);
            }
        }
    }
    double s = 0.0;
    float sf = 0.0f;
    for (int k = 0; k < 8; ++k) {
        s += d[k];
        sf += f[k];
    }
    outd[blockIdx.x * 256 + threadIdx.x] = s;
    outf[blockIdx.x * 256 + threadIdx.x] = sf;
}

template <int OP>
float time_op(double *od, float *of, hipEvent_t e0, hipEvent_t e1) {
    hipLaunchKernelGGL(probe<OP>, dim3(BLOCKS), dim3(256), 0, 0, od, of, 0.999999, 0.999999f, 64);
    (void)hipEventRecord(e0);
    hipLaunchKernelGGL(probe<OP>, dim3(BLOCKS), dim3(256), 0, 0, od, of, 0.999999, 0.999999f,
                       ITERS);
    (void)hipEventRecord(e1);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    return ms;
}

}  // namespace

int main() {
    double *od = nullptr;
    float *of = nullptr;
    hipEvent_t e0, e1;
    if (hipMalloc(&od, (size_t)BLOCKS * 256 * sizeof(double)) != hipSuccess ||
        hipMalloc(&of, (size_t)BLOCKS * 256 * sizeof(float)) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        std::fprintf(stderr, "valu-probe: HIP setup failed\n");
        return 1;
    }
    int cus = 0;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0);
    const char *names[] = {"fma f64",
                           "add f64",
                           "fma f32",
                           "add f32",
                           "cvt f64->f32, cvt f32->f64, mul f64",
                           "rsq f64 + add f64",
                           "rsq f32 + add f32",
                           "fma f64 | fma f32 (independent)",
                           "add f64 -> cvt f64->f32 -> fma f32"};
    const int per_step[] = {1, 1, 1, 1, 3, 2, 2, 2, 3};
    float ms[9];
    ms[0] = time_op<0>(od, of, e0, e1);
    ms[1] = time_op<1>(od, of, e0, e1);
    ms[2] = time_op<2>(od, of, e0, e1);
    ms[3] = time_op<3>(od, of, e0, e1);
    ms[4] = time_op<4>(od, of, e0, e1);
    ms[5] = time_op<5>(od, of, e0, e1);
    ms[6] = time_op<6>(od, of, e0, e1);
    ms[7] = time_op<7>(od, of, e0, e1);
    ms[8] = time_op<8>(od, of, e0, e1);
    const double simds = 4.0 * (cus > 0 ? cus : 256);
    const double steps = (double)BLOCKS * 4 * ITERS * 8 / simds;  // chain steps per SIMD
    std::printf("%-40s %9s %14s %14s\n", "chain step", "ms", "ns/step/SIMD", "ns/instr/SIMD");
    for (int k = 0; k < 9; ++k) {
        double ns = ms[k] * 1e6 / steps;
        std::printf("%-40s %9.3f %14.3f %14.3f\n", names[k], ms[k], ns, ns / per_step[k]);
    }
    (void)hipFree(od);
    (void)hipFree(of);
    return 0;
}
