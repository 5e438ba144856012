// hbm_copy - the calibration kernel for HBM traffic counters: one plain
// device copy of a known number of bytes per launch (16-byte vector loads
// and stores, grid-stride). rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on this
// program gives the counter value of a launch that reads and writes exactly
// `bytes`; the ratio of the two scales the counters read on other kernels.
//
// usage: hbm_copy [MiB per buffer (default 1024)] [launches (default 20)]
// prints one JSON line: bytes read and written per launch, average ms per
// launch (HIP events), GB/s (read + write).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CK(x)                                                                    \
    do {                                                                         \
        hipError_t e_ = (x);                                                     \
        if (e_ != hipSuccess) {                                                  \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));         \
            return 1;                                                            \
        }                                                                        \
    } while (0)

__global__ void __launch_bounds__(256) copy_kernel(const int4* __restrict__ a, int4* __restrict__ b, long n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) b[i] = a[i];
}

__global__ void fill_kernel(int4* __restrict__ a, long n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        a[i] = make_int4((int)i, 1, 2, 3);
}

int main(int argc, char** argv) {
    const long mib = argc > 1 ? std::atol(argv[1]) : 1024;
    const int reps = argc > 2 ? std::atoi(argv[2]) : 20;
    if (mib < 1 || mib > 65536 || reps < 1 || reps > 10000) {
        std::fprintf(stderr, "hbm_copy: MiB in [1, 65536], launches in [1, 10000]\n");
        return 2;
    }
    const long bytes = mib << 20;
    const long n = bytes / (long)sizeof(int4);
    int4 *a = nullptr, *b = nullptr;
    CK(hipMalloc(&a, bytes));
    CK(hipMalloc(&b, bytes));
    const int grid = 256 * 8 * 4;  // 4 blocks of 256 per CU
    fill_kernel<<<grid, 256>>>(a, n);
    fill_kernel<<<grid, 256>>>(b, n);
    copy_kernel<<<grid, 256>>>(a, b, n);
    CK(hipDeviceSynchronize());
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    CK(hipEventRecord(e0));
    for (int r = 0; r < reps; ++r) copy_kernel<<<grid, 256>>>(a, b, n);
    CK(hipEventRecord(e1));
    CK(hipEventSynchronize(e1));
    float ms = 0;
    CK(hipEventElapsedTime(&ms, e0, e1));
    CK(hipGetLastError());
    const double per = ms / reps;
    std::printf("{\"workload\": \"hbm_copy\", \"kernel\": \"copy_kernel\", \"bytes_read\": %ld, \"bytes_written\": %ld, "
                "\"launches\": %d, \"ms_per_launch\": %.6f, \"GBps\": %.1f}\n",
                bytes, bytes, reps + 1, per, 2.0 * bytes / (per * 1e-3) / 1e9);
    CK(hipFree(a));
    CK(hipFree(b));
    return 0;
}
