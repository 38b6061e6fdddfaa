// Batch ray queries on a triangle mesh: the C-ABI owe_hip_mesh_* and owe_hip_trace_batch.
// Stands in for MeshShape::intersect over many rays (src/owe/scene/mesh.cpp:73-97) and the float
// kernels' any-hit occludedMesh (src/owe/backends/gpu/shaders/geometry.slang:577-615).
//
// Two launch forms. One lane per ray (traceBatchKernel), 128 lanes per block: a wave lives until its
// longest ray ends. Lane refill (traceRefillKernel): a persistent grid whose waves each own a
// contiguous range of rays; a lane whose ray ends waits until `refill` lanes of its wave are idle,
// then the idle lanes take the next rays of the range together, so the wave keeps most of its lanes
// busy instead of waiting on its slowest ray. Both forms walk each ray identically and return the
// same hits bit for bit. The traversal stack lives in LDS with the capacity the build proved
// sufficient (bvh4.hpp), chosen from a few compiled sizes.
#include <hip/hip_runtime.h>

#include <string>

#include "bvh4.hpp"
#include "common.hpp"
#include "traverse.hpp"

using namespace owehip;

struct owe_hip_mesh {
    int device = 0;
    WideNode* nodes = nullptr;
    TriRecord* tris = nullptr;
    owe_hip_mesh_stats stats{};
    int cus = 0;                 // compute units of the device
    // Per query mode (closest, any): the launch form, and the refill form's resident blocks per
    // compute unit and idle lanes that trigger a refill.
    int form[2] = {OWE_HIP_TRACE_REFILL, OWE_HIP_TRACE_REFILL};
    int blocksPerCu[2] = {8, 8};
    int refill[2] = {32, 32};
};

namespace owehip {
namespace {

constexpr int kQueryBlock = 128;

template <int kCap, bool kAny>
__global__ __launch_bounds__(kQueryBlock) void traceBatchKernel(const WideNode* __restrict__ nodes,
                                                                const TriRecord* __restrict__ tris,
                                                                const float4* __restrict__ rays, uint64_t n,
                                                                uint32_t* __restrict__ primOut,
                                                                float* __restrict__ tOut) {
    __shared__ uint32_t stack[kCap * kQueryBlock];
    const uint64_t i = uint64_t(blockIdx.x) * kQueryBlock + threadIdx.x;
    if (i >= n) return;  // no barrier follows: lanes past the end simply leave
    const float4 r0 = rays[2 * i], r1 = rays[2 * i + 1];
    const RayHit h = traverseBvh4<kAny>(nodes, tris, mk3(r0.x, r0.y, r0.z), mk3(r1.x, r1.y, r1.z), r0.w, r1.w, ~0u,
                                        stack + threadIdx.x, kQueryBlock);
    if (h.slot == kEmptyChild) {
        primOut[i] = 0xFFFFFFFFu;
        tOut[i] = __builtin_inff();
    } else {
        primOut[i] = tris[h.slot].prim;
        tOut[i] = h.t;
    }
}

// Result staging of the refill form. A wave takes its rays in increasing order, so they end in nearly
// that order, but a lane's store of its own result touches a cache line that no neighbour has filled
// yet. Each wave therefore holds kStageSlots LDS slots of 64 results: group g of its range (rays
// start + 64g … start + 64g + 63) is staged in slot g % kStageSlots when that slot is free as the
// wave first takes a ray of g, and leaves as one coalesced 64-lane store per output array once all
// of its rays have ended. A group whose slot is still held writes its results directly, so a slow
// ray never holds back the refill.
constexpr uint32_t kStageSlots = 4;
constexpr uint32_t kNoGroup = 0xFFFFFFFFu;

// Persistent lanes with refill: wave w owns rays [w·per, min(n, (w + 1)·per)). Every branch that
// decides a refill or a flush is wave-uniform (ballots, the range and the slots' owners); no barrier
// is used, so waves run independently and each exits once its range is drained and its lanes are
// idle, by which time every staged group has been flushed.
template <int kCap, bool kAny>
__global__ __launch_bounds__(kQueryBlock) void traceRefillKernel(const WideNode* __restrict__ nodes,
                                                                 const TriRecord* __restrict__ tris,
                                                                 const float4* __restrict__ rays, uint64_t n,
                                                                 uint64_t per, uint32_t refill,
                                                                 uint32_t* __restrict__ primOut,
                                                                 float* __restrict__ tOut) {
    __shared__ uint32_t stack[kCap * kQueryBlock];
    __shared__ uint2 staged[kQueryBlock / 64][kStageSlots][64];
    const uint32_t lane = threadIdx.x & 63u;
    uint2(*stage)[64] = staged[threadIdx.x >> 6];
    const uint64_t wave = (uint64_t(blockIdx.x) * kQueryBlock + threadIdx.x) >> 6;
    const uint64_t start = wave * per;
    uint64_t next = start;
    const uint64_t end = next < n ? (n - next < per ? n : next + per) : next;
    uint32_t owner[kStageSlots], held[kStageSlots];
#pragma unroll
    for (uint32_t j = 0; j < kStageSlots; ++j) owner[j] = kNoGroup, held[j] = 0;
    uint32_t opened = 0;  // groups of the range whose slot has been decided
    TravState s;
    uint64_t ray = 0;
    bool active = false;
    for (;;) {
        const uint64_t idle = __ballot(!active);
        const uint32_t nIdle = uint32_t(__popcll(idle));
        if (next < end && (nIdle >= refill || nIdle == 64u)) {
            const uint32_t rank = uint32_t(__popcll(idle & ((1ull << lane) - 1ull)));
            if (!active && next + rank < end) {
                ray = next + rank;
                const float4 r0 = rays[2 * ray], r1 = rays[2 * ray + 1];
                beginBvh4(s, mk3(r0.x, r0.y, r0.z), mk3(r1.x, r1.y, r1.z), r0.w, r1.w);
                active = true;
            }
            next = end - next > nIdle ? next + nIdle : end;
            // The taken rays (at least one: refill ≥ 1) span at most two new groups; each takes its
            // slot if the slot is free.
            const uint32_t last = uint32_t((next - 1 - start) >> 6);
            for (; opened <= last; ++opened) {
#pragma unroll
                for (uint32_t j = 0; j < kStageSlots; ++j)
                    if (opened % kStageSlots == j && owner[j] == kNoGroup) owner[j] = opened, held[j] = 0;
            }
        }
        if (__ballot(active) == 0) break;
        bool done = false;
        uint32_t prim = 0;
        float t = 0.0f;
        if (active && stepBvh4<kAny>(s, nodes, tris, stack + threadIdx.x, kQueryBlock)) {
            prim = s.hit.slot == kEmptyChild ? 0xFFFFFFFFu : tris[s.hit.slot].prim;
            t = s.hit.slot == kEmptyChild ? __builtin_inff() : s.hit.t;
            active = false;
            done = true;
        }
        if (__ballot(done) == 0) continue;
        const uint32_t group = uint32_t((ray - start) >> 6), slot = group % kStageSlots;
        bool staging = false;
#pragma unroll
        for (uint32_t j = 0; j < kStageSlots; ++j) staging |= slot == j && owner[j] == group;
        if (done) {
            if (staging) {
                stage[slot][(ray - start) & 63u] = make_uint2(prim, __float_as_uint(t));
            } else {
                primOut[ray] = prim;
                tOut[ray] = t;
            }
        }
#pragma unroll
        for (uint32_t j = 0; j < kStageSlots; ++j) {
            held[j] += uint32_t(__popcll(__ballot(done && staging && slot == j)));
            if (owner[j] == kNoGroup) continue;
            const uint64_t base = start + 64ull * owner[j];
            const uint32_t size = end - base < 64u ? uint32_t(end - base) : 64u;
            if (held[j] != size) continue;
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            if (lane < size) {
                const uint2 e = stage[j][lane];
                primOut[base + lane] = e.x;
                tOut[base + lane] = __uint_as_float(e.y);
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            owner[j] = kNoGroup;
            held[j] = 0;
        }
    }
}

template <int kCap>
void launchTrace(int mode, const owe_hip_mesh* m, const float* rays, uint64_t n, uint32_t* prim, float* t,
                 hipStream_t s) {
    const float4* r4 = reinterpret_cast<const float4*>(rays);
    const int q = mode == OWE_HIP_QUERY_ANY ? 1 : 0;
    if (m->form[q] == OWE_HIP_TRACE_REFILL) {
        const uint64_t blocks = uint64_t(m->cus) * uint64_t(m->blocksPerCu[q]);
        const uint64_t waves = blocks * (kQueryBlock / 64);
        const uint64_t per = (n + waves - 1) / waves;
        const dim3 grid{unsigned(blocks)}, block{unsigned(kQueryBlock)};
        if (mode == OWE_HIP_QUERY_ANY)
            hipLaunchKernelGGL((traceRefillKernel<kCap, true>), grid, block, 0, s, m->nodes, m->tris, r4, n, per,
                               uint32_t(m->refill[q]), prim, t);
        else
            hipLaunchKernelGGL((traceRefillKernel<kCap, false>), grid, block, 0, s, m->nodes, m->tris, r4, n, per,
                               uint32_t(m->refill[q]), prim, t);
        OWE_HIP_CHECK(hipGetLastError());
        return;
    }
    const uint64_t blocks = (n + kQueryBlock - 1) / kQueryBlock;
    if (blocks > 0x7FFFFFFFull) fail(OWE_HIP_ERR_INVALID, "batch too large for one launch");
    const dim3 grid{unsigned(blocks)}, block{unsigned(kQueryBlock)};
    if (mode == OWE_HIP_QUERY_ANY)
        hipLaunchKernelGGL((traceBatchKernel<kCap, true>), grid, block, 0, s, m->nodes, m->tris,
                           reinterpret_cast<const float4*>(rays), n, prim, t);
    else
        hipLaunchKernelGGL((traceBatchKernel<kCap, false>), grid, block, 0, s, m->nodes, m->tris,
                           reinterpret_cast<const float4*>(rays), n, prim, t);
    OWE_HIP_CHECK(hipGetLastError());
}

}  // namespace
}  // namespace owehip

extern "C" {

int owe_hip_mesh_create(int device, const float* positions, uint32_t nVertices, const uint32_t* triangles,
                        uint32_t nTriangles, owe_hip_mesh** out) {
    return guard([&] {
        require(out != nullptr, "out is null");
        require(positions != nullptr || nVertices == 0, "positions is null");
        require(triangles != nullptr || nTriangles == 0, "triangles is null");
        *out = nullptr;
        selectDevice(device);
        Bvh4 b = buildBvh4(positions, nVertices, triangles, nTriangles);
        if (b.stackBound > 64)
            fail(OWE_HIP_ERR_UNSUPPORTED, "mesh hierarchy needs a traversal stack of " + std::to_string(b.stackBound) +
                                              " entries; the kernels hold at most 64");
        auto* m = new owe_hip_mesh;
        m->device = device;
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) != hipSuccess || prop.multiProcessorCount <= 0) {
            delete m;
            fail(OWE_HIP_ERR_DEVICE, "cannot read the device's compute-unit count");
        }
        m->cus = prop.multiProcessorCount;
        const size_t nb = b.nodes.size() * sizeof(WideNode), tb = std::max<size_t>(1, b.tris.size()) * sizeof(TriRecord);
        hipError_t e = hipMalloc(&m->nodes, nb);
        if (e == hipSuccess) e = hipMalloc(&m->tris, tb);
        if (e == hipSuccess) e = hipMemcpy(m->nodes, b.nodes.data(), nb, hipMemcpyHostToDevice);
        if (e == hipSuccess && !b.tris.empty())
            e = hipMemcpy(m->tris, b.tris.data(), b.tris.size() * sizeof(TriRecord), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            if (m->nodes) (void)hipFree(m->nodes);
            if (m->tris) (void)hipFree(m->tris);
            delete m;
            hipCheck(e, "mesh upload");
        }
        m->stats.triangles = nTriangles;
        m->stats.nodes = b.nodes.size();
        m->stats.depth = b.depth;
        m->stats.stack_bound = b.stackBound;
        m->stats.build_seconds = b.seconds;
        m->stats.device_bytes = nb + tb;
        *out = m;
    });
}

int owe_hip_mesh_destroy(owe_hip_mesh* m) {
    return guard([&] {
        if (!m) return;
        (void)hipSetDevice(m->device);
        if (m->nodes) OWE_HIP_CHECK(hipFree(m->nodes));
        if (m->tris) OWE_HIP_CHECK(hipFree(m->tris));
        delete m;
    });
}

int owe_hip_mesh_stats_get(const owe_hip_mesh* m, owe_hip_mesh_stats* out) {
    return guard([&] {
        require(m != nullptr && out != nullptr, "null argument");
        *out = m->stats;
    });
}

int owe_hip_mesh_set_launch(owe_hip_mesh* m, int mode, int form, int blocks_per_cu, int refill) {
    return guard([&] {
        require(m != nullptr, "mesh is null");
        require(mode == -1 || mode == OWE_HIP_QUERY_CLOSEST || mode == OWE_HIP_QUERY_ANY,
                "mode must be closest (0), any (1) or both (-1)");
        require(form == OWE_HIP_TRACE_PER_LANE || form == OWE_HIP_TRACE_REFILL, "form must be 0 (per lane) or 1 (refill)");
        require(blocks_per_cu >= 0 && blocks_per_cu <= 10, "blocks_per_cu must be 0 (keep) to 10");
        require(refill >= 0 && refill <= 64, "refill must be 0 (keep) to 64");
        for (int q = 0; q < 2; ++q) {
            if (mode != -1 && mode != q) continue;
            m->form[q] = form;
            if (blocks_per_cu) m->blocksPerCu[q] = blocks_per_cu;
            if (refill) m->refill[q] = refill;
        }
    });
}

int owe_hip_mesh_get_launch(const owe_hip_mesh* m, int mode, int* form, int* blocks_per_cu, int* refill) {
    return guard([&] {
        require(m != nullptr && form != nullptr && blocks_per_cu != nullptr && refill != nullptr, "null argument");
        require(mode == OWE_HIP_QUERY_CLOSEST || mode == OWE_HIP_QUERY_ANY, "mode must be closest (0) or any (1)");
        *form = m->form[mode];
        *blocks_per_cu = m->blocksPerCu[mode];
        *refill = m->refill[mode];
    });
}

int owe_hip_trace_batch(owe_hip_mesh* m, int mode, const float* rays, uint64_t n, uint32_t* prim, float* t,
                        void* stream) {
    return guard([&] {
        require(m != nullptr, "mesh is null");
        require(mode == OWE_HIP_QUERY_CLOSEST || mode == OWE_HIP_QUERY_ANY, "mode must be closest (0) or any (1)");
        if (n == 0) return;
        require(rays != nullptr && prim != nullptr && t != nullptr, "null buffer");
        require((reinterpret_cast<uintptr_t>(rays) & 15) == 0, "rays must be 16-byte aligned");
        for (const void* p : {static_cast<const void*>(rays), static_cast<const void*>(prim), static_cast<const void*>(t)}) {
            hipPointerAttribute_t a;
            if (hipPointerGetAttributes(&a, p) != hipSuccess || a.type != hipMemoryTypeDevice || a.device != m->device)
                fail(OWE_HIP_ERR_INVALID, "rays, prim and t must be device memory on the mesh's device");
        }
        OWE_HIP_CHECK(hipSetDevice(m->device));
        const hipStream_t s = static_cast<hipStream_t>(stream);
        const uint32_t cap = m->stats.stack_bound;
        if (cap <= 16) launchTrace<16>(mode, m, rays, n, prim, t, s);
        else if (cap <= 32) launchTrace<32>(mode, m, rays, n, prim, t, s);
        else if (cap <= 48) launchTrace<48>(mode, m, rays, n, prim, t, s);
        else launchTrace<64>(mode, m, rays, n, prim, t, s);
    });
}

}  // extern "C"
