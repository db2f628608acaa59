// Vector-memory return-rate probe: measurement code of the benchmark (measure/libchunk_probe.so),
// not part of the library or of its drop-in boundary.
//
// A traversal lane reads its own record: 14 consecutive 16-byte chunks of a 256-byte node, or 5
// of an 80-byte leaf record (two triangles), at addresses no other lane shares. What bounds such a kernel is the
// rate at which a compute unit returns 16-byte chunks to lanes, not HBM bandwidth. This kernel
// reproduces that access pattern without the traversal: every lane reads `chunks` consecutive
// chunks of a pseudo-random 256-byte record, `iters` times, and the host divides chunks by time.
#include <hip/hip_runtime.h>
#include <stdint.h>


namespace {

__device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

template <int K>
__global__ __launch_bounds__(256) void k_chunk_probe(const float4* __restrict__ buf, uint32_t n_rec, int iters,
                                                     float* __restrict__ out) {
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    float acc = 0.f;
    uint32_t h = mix32(gid * 2654435761u + 0x5eedu);
    for (int i = 0; i < iters; i++) {
        h = mix32(h + (uint32_t)i);
        const float4* rec = buf + (size_t)(h % n_rec) * 16;  // record index < n_rec: in bounds
        // all K loads issue back to back, as a traversal lane issues a node's or a triangle's
        float4 v[K];
#pragma unroll
        for (int k = 0; k < K; k++) v[k] = rec[k];
#pragma unroll
        // every component is used, so each chunk stays one 16-byte load (a sum of fewer components
        // lets the compiler load only those, as 4- and 8-byte loads straddling the chunks)
        for (int k = 0; k < K; k++) acc += (v[k].x + v[k].y) + (v[k].z + v[k].w);
    }
    out[gid] = acc;  // keeps the loads live
}

template <int K>
hipError_t launch(const void* buf, uint32_t n_rec, int iters, int blocks, float* out, hipStream_t s) {
    hipLaunchKernelGGL(k_chunk_probe<K>, dim3((unsigned)blocks), dim3(256), 0, s, static_cast<const float4*>(buf), n_rec,
                       iters, out);
    return hipGetLastError();
}

}  // namespace

// Launches `blocks` workgroups of 256 lanes on `stream` (a hipStream_t, NULL = default stream).
// Every lane reads, `iters` times, `chunks` (1, 3, 5, 7, 8, 14 or 16) consecutive 16-byte chunks of a
// pseudo-random 256-byte record of `buf` (a device buffer of n_rec * 256 bytes), all chunks of a
// record issued back to back, and writes one float to out[lane] (blocks * 256 floats). Returns 0
// or a hipError_t code (other chunk counts are refused); asynchronous.
extern "C" int chunk_probe(const void* buf, uint32_t n_rec, int chunks, int iters, int blocks, float* out,
                           void* stream);

extern "C" int chunk_probe(const void* buf, uint32_t n_rec, int chunks, int iters, int blocks, float* out,
                           void* stream) {
    if (!buf || !out || n_rec == 0 || iters < 1 || blocks < 1) return (int)hipErrorInvalidValue;
    const hipStream_t s = static_cast<hipStream_t>(stream);
    switch (chunks) {
        case 1: return (int)launch<1>(buf, n_rec, iters, blocks, out, s);
        case 3: return (int)launch<3>(buf, n_rec, iters, blocks, out, s);
        case 5: return (int)launch<5>(buf, n_rec, iters, blocks, out, s);
        case 7: return (int)launch<7>(buf, n_rec, iters, blocks, out, s);
        case 8: return (int)launch<8>(buf, n_rec, iters, blocks, out, s);
        case 14: return (int)launch<14>(buf, n_rec, iters, blocks, out, s);
        case 16: return (int)launch<16>(buf, n_rec, iters, blocks, out, s);
        default: return (int)hipErrorInvalidValue;
    }
}
