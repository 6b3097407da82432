// Level-synchronisation cost on one GPU: what separates two BFS levels.
//
//   boundary  N back-to-back launches of an empty kernel (grid of `blocks` x 512) on one
//             stream: mean time per launch = the cost of a kernel boundary;
//   replayed  the same N launches captured once in a HIP graph and replayed;
//   barrier   one cooperative launch (blocks no larger than the device can hold at
//             once, checked with the occupancy query before launching) that passes N
//             grid-wide barriers: mean time per barrier.
// The traversal kernels are designed from these numbers: a level is two stream-ordered
// kernels, not a grid barrier inside one persistent kernel.
//
// Usage: level_sync_probe [N=2000]; prints one line per block count.
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

namespace cg = cooperative_groups;

#define CHECK(x)                                                              \
    do {                                                                      \
        hipError_t e_ = (x);                                                  \
        if (e_ != hipSuccess) {                                               \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));     \
            std::exit(1);                                                     \
        }                                                                     \
    } while (0)

__global__ void empty_kernel(int* sink) {
    if (sink && threadIdx.x == 0) *sink = 1;  // never taken: sink is null
}

__global__ void barrier_kernel(int rounds, unsigned* counter) {
    cg::grid_group grid = cg::this_grid();
    for (int i = 0; i < rounds; ++i) {
        if (threadIdx.x == 0) atomicAdd(counter, 1u);
        grid.sync();
    }
}

int main(int argc, char** argv) {
    const int N = argc > 1 ? std::atoi(argv[1]) : 2000;
    hipDeviceProp_t prop;
    CHECK(hipGetDeviceProperties(&prop, 0));
    const int cus = prop.multiProcessorCount;
    hipStream_t s;
    CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    hipEvent_t a, b;
    CHECK(hipEventCreate(&a));
    CHECK(hipEventCreate(&b));
    unsigned* counter;
    CHECK(hipMalloc(&counter, sizeof(unsigned)));
    int perCu = 0;
    CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, barrier_kernel, 512, 0));
    std::printf("device %s, %d CUs, barrier kernel: %d resident workgroups of 512 per CU\n", prop.name, cus, perCu);
    for (int per : {1, 2, 4}) {
        const int blocks = cus * per;
        // kernel boundaries
        for (int w = 0; w < 50; ++w) hipLaunchKernelGGL(empty_kernel, dim3(blocks), dim3(512), 0, s, nullptr);
        CHECK(hipEventRecord(a, s));
        for (int i = 0; i < N; ++i) hipLaunchKernelGGL(empty_kernel, dim3(blocks), dim3(512), 0, s, nullptr);
        CHECK(hipEventRecord(b, s));
        CHECK(hipEventSynchronize(b));
        float msBoundary = 0;
        CHECK(hipEventElapsedTime(&msBoundary, a, b));
        // the same launches replayed from a captured graph
        hipGraph_t graph;
        hipGraphExec_t exec;
        CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        for (int i = 0; i < N; ++i) hipLaunchKernelGGL(empty_kernel, dim3(blocks), dim3(512), 0, s, nullptr);
        CHECK(hipStreamEndCapture(s, &graph));
        CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        CHECK(hipGraphLaunch(exec, s));
        CHECK(hipEventRecord(a, s));
        CHECK(hipGraphLaunch(exec, s));
        CHECK(hipEventRecord(b, s));
        CHECK(hipEventSynchronize(b));
        float msGraph = 0;
        CHECK(hipEventElapsedTime(&msGraph, a, b));
        CHECK(hipGraphExecDestroy(exec));
        CHECK(hipGraphDestroy(graph));
        // grid barriers (only when the whole grid can be resident)
        float msBarrier = -1;
        if (per <= perCu) {
            int rounds = N;
            void* args[] = {&rounds, &counter};
            unsigned zero = 0;
            CHECK(hipMemcpy(counter, &zero, sizeof(unsigned), hipMemcpyHostToDevice));
            CHECK(hipEventRecord(a, s));
            CHECK(hipLaunchCooperativeKernel(reinterpret_cast<void*>(barrier_kernel), dim3(blocks), dim3(512), args, 0,
                                             s));
            CHECK(hipEventRecord(b, s));
            CHECK(hipEventSynchronize(b));
            CHECK(hipEventElapsedTime(&msBarrier, a, b));
            unsigned got = 0;
            CHECK(hipMemcpy(&got, counter, sizeof(unsigned), hipMemcpyDeviceToHost));
            if (got != unsigned(rounds) * unsigned(blocks)) std::printf("barrier count mismatch: %u\n", got);
        }
        if (msBarrier < 0)
            std::printf("workgroups %5d: kernel boundary %.2f us, replayed from a graph %.2f us, grid barrier: "
                        "grid not resident\n",
                        blocks, msBoundary * 1e3f / N, msGraph * 1e3f / N);
        else
            std::printf("workgroups %5d: kernel boundary %.2f us, replayed from a graph %.2f us, grid barrier %.2f us\n",
                        blocks, msBoundary * 1e3f / N, msGraph * 1e3f / N, msBarrier * 1e3f / N);
    }
    return 0;
}
