// valu_probe.hip -- issue cost of the VALU instruction kinds the pair
// kernels are made of, on the current GPU.  Each kernel runs 8 independent
// dependency chains per lane, 2048 workgroups of 256 threads (8 waves per
// SIMD on a 256-CU part), all in registers; the report is nanoseconds per
// chain step per SIMD, i.e. per wave64 instruction group of the listed
// kind.  Second part: the relative error of the v_rsq_f64 and v_rsq_f32
// estimates that seed the pair kernels' 1/r (one Newton step has to take
// them to the FP64 path's accuracy).  Used to explain and bound the pair
// kernels; not part of the library.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

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

__global__ __launch_bounds__(256) void rsq_seeds(const double *x, double *y, float *yf, int n) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        y[i] = __builtin_amdgcn_rsq(x[i]);
        yf[i] = __builtin_amdgcn_rsqf((float)x[i]);
    }
}

// Largest |y sqrt(x) - 1| of the hardware estimates over 2^22 inputs: a
// uniform sweep of [1, 4) (the estimate's period in x is a factor of 4)
// and log-uniform values over 2^-60 .. 2^60; the reference is long double.
int rsq_accuracy() {
    const int n = 1 << 22;
    std::vector<double> x(n), y(n);
    std::vector<float> yf(n);
    std::mt19937_64 g(12345);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    for (int i = 0; i < n / 2; ++i) x[i] = 1.0 + 3.0 * (i + 0.5) / (n / 2);
    for (int i = n / 2; i < n; ++i) x[i] = std::ldexp(1.0 + u(g), (int)(u(g) * 120) - 60);
    double *dx = nullptr, *dy = nullptr;
    float *dyf = nullptr;
    if (hipMalloc(&dx, n * sizeof(double)) != hipSuccess ||
        hipMalloc(&dy, n * sizeof(double)) != hipSuccess ||
        hipMalloc(&dyf, n * sizeof(float)) != hipSuccess ||
        hipMemcpy(dx, x.data(), n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        std::fprintf(stderr, "valu-probe: HIP setup failed\n");
        return 1;
    }
    hipLaunchKernelGGL(rsq_seeds, dim3(n / 256), dim3(256), 0, 0, dx, dy, dyf, n);
    if (hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(y.data(), dy, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(yf.data(), dyf, n * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) {
        std::fprintf(stderr, "valu-probe: rsq run failed\n");
        return 1;
    }
    long double e64 = 0, e32 = 0;
    for (int i = 0; i < n; ++i) {
        const long double xd = x[i], xf = (float)x[i];
        e64 = std::fmax(e64, std::fabs((long double)y[i] * std::sqrt(xd) - 1.0L));
        e32 = std::fmax(e32, std::fabs((long double)yf[i] * std::sqrt(xf) - 1.0L));
    }
    std::printf("rsq estimate, max relative error over %d inputs: v_rsq_f64 %.3Le (2^%.2f), "
                "v_rsq_f32 %.3Le (2^%.2f)\n",
                n, e64, (double)std::log2(e64), e32, (double)std::log2(e32));
    (void)hipFree(dx);
    (void)hipFree(dy);
    (void)hipFree(dyf);
    return 0;
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
    return rsq_accuracy();
}
