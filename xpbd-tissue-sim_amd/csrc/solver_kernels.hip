// XPBD solver kernels for gfx950 (MI355X).
//
// Vertex state is one 32-byte record per vertex: float4 base (displacement at
// the start of the step, inverse mass in w) followed by float4 delta (motion in
// the current step), so a tet's four random-access gathers touch four aligned
// 32-byte blocks.  A separate float4 per vertex keeps the low-order part of the
// base (compensated once per step) and is read only outside the solve.
// Element data is structure-of-arrays in solver order so consecutive lanes
// read consecutive words.  All kernels use 64-lane wavefronts and
// 256-thread workgroups.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "mesh_tools.h"
#include "xpbd_math.h"

namespace xpbd {

namespace {

constexpr int kBlock = 256;

struct Gather {
    float b[12];  // base displacement of the 4 vertices
    float d[12];  // in-step motion
    float w[4];   // inverse masses
};

__device__ __forceinline__ void gather4(const float4* __restrict__ pos, const int4 t, Gather& g)
{
    const int vid[4] = { t.x, t.y, t.z, t.w };
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const float4 b = pos[2 * (size_t)vid[i]];
        const float4 d = pos[2 * (size_t)vid[i] + 1];
        g.b[3 * i] = b.x;
        g.b[3 * i + 1] = b.y;
        g.b[3 * i + 2] = b.z;
        g.w[i] = b.w;
        g.d[3 * i] = d.x;
        g.d[3 * i + 1] = d.y;
        g.d[3 * i + 2] = d.z;
    }
}

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b runs on
// XCD b % 8), and each XCD has its own L2.  With remap set, workgroup b takes
// the work item that gives every XCD one contiguous eighth of the G items:
// neighbouring clusters or patches (adjacent in solver order) then share an L2,
// and since every colour's clusters follow the same spatial order, a vertex
// written in one colour launch is read in the next by the same XCD.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_item(int b, int G, bool remap)
{
    if (!remap) return b;
    const int q = G / kXcds, r = G % kXcds, x = b % kXcds;
    return x * q + min(x, r) + b / kXcds;
}

__device__ __forceinline__ void load_tet(const TetArrays& ta, int e, float* q, float2& at, float& g)
{
#pragma unroll
    for (int k = 0; k < 9; k++) q[k] = ta.Q[(size_t)k * ta.qstride + e];
    at = ta.at[e];
    g = ta.gamma[e];
}

// ---------------------------------------------------------------------------
// predict: x_prev = x ; x += dt v ; z += -g dt^2   (XPBDMeshObject.cpp:426-470)
// The step's motion starts as delta = dt v + (0, 0, -g dt^2); the base is the
// position at the start of the step, so x_prev needs no copy.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_predict(float4* __restrict__ pos, const float4* __restrict__ vel,
                                                   int nv, float dt, float dz)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= nv) return;
    const float4 v = vel[i];
    pos[2 * (size_t)i + 1] = make_float4(dt * v.x, dt * v.y, dt * v.z + dz, 0.f);
}

// first-order inertial move (XPBDMeshObject.cpp:450-457): z -= g m_i dt / B_i,
// no velocity term
__global__ __launch_bounds__(kBlock) void k_predict_first_order(float4* __restrict__ pos,
                                                               const float* __restrict__ dz1, int nv)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= nv) return;
    pos[2 * (size_t)i + 1] = make_float4(0.f, 0.f, dz1[i], 0.f);
}

// ---------------------------------------------------------------------------
// finish: pinned vertices back to x_prev (XPBDMeshObject.cpp:497-503), then
// v = (x - x_prev) / dt (XPBDMeshObject.cpp:507-535) = delta / dt, and the
// base absorbs delta with a compensated (two-sum) update.  vel.w != 0 marks pins.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e)
{
    s = a + b;
    const float bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

__global__ __launch_bounds__(kBlock) void k_finish(float4* __restrict__ pos, float4* __restrict__ lo,
                                                  float4* __restrict__ vel, int nv, float dt,
                                                  unsigned long long* __restrict__ step)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i == 0) step[0] = step[0] + 1ull;
    if (i >= nv) return;
    float4 b = pos[2 * (size_t)i];
    float4 d = pos[2 * (size_t)i + 1];
    float4 l = lo[i];
    float4 v = vel[i];
    if (v.w != 0.0f) d = make_float4(0.f, 0.f, 0.f, 0.f);
    v.x = d.x / dt;
    v.y = d.y / dt;
    v.z = d.z / dt;
    vel[i] = v;
    two_sum(b.x, d.x + l.x, b.x, l.x);
    two_sum(b.y, d.y + l.y, b.y, l.y);
    two_sum(b.z, d.z + l.z, b.z, l.z);
    pos[2 * (size_t)i] = b;
    pos[2 * (size_t)i + 1] = make_float4(0.f, 0.f, 0.f, 0.f);
    lo[i] = l;
}

// The in-step motion folded into the base (compensated, as k_finish does),
// without the velocity update, the pin reset or the step counter: positions
// moved between steps (a subset solve) stay where they are.
__global__ __launch_bounds__(kBlock) void k_commit(float4* __restrict__ pos, float4* __restrict__ lo, int nv)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= nv) return;
    float4 b = pos[2 * (size_t)i];
    const float4 d = pos[2 * (size_t)i + 1];
    float4 l = lo[i];
    two_sum(b.x, d.x + l.x, b.x, l.x);
    two_sum(b.y, d.y + l.y, b.y, l.y);
    two_sum(b.z, d.z + l.z, b.z, l.z);
    pos[2 * (size_t)i] = b;
    pos[2 * (size_t)i + 1] = make_float4(0.f, 0.f, 0.f, 0.f);
    lo[i] = l;
}

// Multipliers of the listed elements (mult > 0) back to zero: a projector's
// initialize() (XPBDSolver.hpp:108-116).
__global__ __launch_bounds__(kBlock) void k_reset_listed(float2* __restrict__ lam, const int* __restrict__ mult, int nt)
{
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e < nt && mult[e] > 0) lam[e] = make_float2(0.f, 0.f);
}

// ---------------------------------------------------------------------------
// Colour-ordered Gauss-Seidel: one launch per colour; tets of one colour share
// no vertex, so each thread reads and writes its four vertices directly.
// ---------------------------------------------------------------------------
template <bool FIRST>
__global__ __launch_bounds__(kBlock) void k_gs_color(TetArrays ta, float4* __restrict__ pos, int begin, int end)
{
    const int e = begin + blockIdx.x * kBlock + threadIdx.x;
    if (e >= end) return;
    const int4 t = ta.tet[e];
    Gather g;
    gather4(pos, t, g);
    float q[9];
    float2 at;
    float gm;
    load_tet(ta, e, q, at, gm);
    float l[2] = { 0.0f, 0.0f };
    if (!FIRST) {
        const float2 lv = ta.lam[e];
        l[0] = lv.x;
        l[1] = lv.y;
    }
    float dx[12];
    project_tet_split<float>(g.b, g.d, g.w, q, at.x, at.y, gm, l, dx);
    ta.lam[e] = make_float2(l[0], l[1]);
    const int vid[4] = { t.x, t.y, t.z, t.w };
#pragma unroll
    for (int i = 0; i < 4; i++)
        pos[2 * (size_t)vid[i] + 1] =
            make_float4(g.d[3 * i] + dx[3 * i], g.d[3 * i + 1] + dx[3 * i + 1], g.d[3 * i + 2] + dx[3 * i + 2], 0.f);
}

// ---------------------------------------------------------------------------
// Cluster Gauss-Seidel: one thread per cluster (up to 8 vertices, up to 8
// tets, consecutive in solver order).  Clusters of one colour share no vertex.
// The workgroup first stages the element data of its clusters (one contiguous
// tet range) into LDS with coalesced loads; each thread stages its cluster's
// vertices in its own LDS column (slot-major, so lane t always hits bank t),
// projects the tets in order, and writes the vertices back: sequential
// Gauss-Seidel in (colour, cluster, tet) order.  After the one staging
// barrier no thread reads another thread's vertex column.
// ---------------------------------------------------------------------------
constexpr int kClusterBlock = 64;
constexpr int kClusterStage = kClusterBlock * 8;  // at most 8 tets per cluster

template <bool FIRST>
__global__ __launch_bounds__(kClusterBlock) void k_gs_cluster(TetArrays ta, const int4* __restrict__ cverts,
                                                             const int* __restrict__ cstart,
                                                             const uint32_t* __restrict__ tloc,
                                                             float4* __restrict__ pos, int cbeg, int cend, int xcd)
{
    __shared__ float sb[8][4][kClusterBlock];   // base xyz, inverse mass (thread-private columns)
    __shared__ float sd[8][3][kClusterBlock];   // in-step motion
    __shared__ float sq[12][kClusterStage];     // Q (9), alpha tilde (2), gamma of the block's tets
    __shared__ float sl[2][kClusterStage];      // lambda
    __shared__ uint32_t stl[kClusterStage];     // local vertex slots
    const int tid = threadIdx.x;
    const int cb = cbeg + xcd_item(blockIdx.x, gridDim.x, xcd != 0) * kClusterBlock;
    const int cl_end = min(cend, cb + kClusterBlock);
    // the block's clusters are consecutive, so their tets form one contiguous
    // range: stage its element data with coalesced loads
    const int tb = cstart[cb];
    const int n = cstart[cl_end] - tb;
    for (int i = tid; i < n; i += kClusterBlock) {
#pragma unroll
        for (int k = 0; k < 9; k++) sq[k][i] = ta.Q[(size_t)k * ta.qstride + tb + i];
        const float2 a = ta.at[tb + i];
        sq[9][i] = a.x;
        sq[10][i] = a.y;
        sq[11][i] = ta.gamma[tb + i];
        if (!FIRST) {
            const float2 lv = ta.lam[tb + i];
            sl[0][i] = lv.x;
            sl[1][i] = lv.y;
        }
        stl[i] = tloc[tb + i];
    }
    const int c = cb + tid;
    const bool active = c < cend;
    int vid[8];
    if (active) {
        const int4 va = cverts[2 * (size_t)c], vb = cverts[2 * (size_t)c + 1];
        vid[0] = va.x; vid[1] = va.y; vid[2] = va.z; vid[3] = va.w;
        vid[4] = vb.x; vid[5] = vb.y; vid[6] = vb.z; vid[7] = vb.w;
#pragma unroll
        for (int q = 0; q < 8; q++) {
            if (vid[q] < 0) continue;
            const float4 b = pos[2 * (size_t)vid[q]];
            const float4 d = pos[2 * (size_t)vid[q] + 1];
            sb[q][0][tid] = b.x;
            sb[q][1][tid] = b.y;
            sb[q][2][tid] = b.z;
            sb[q][3][tid] = b.w;
            sd[q][0][tid] = d.x;
            sd[q][1][tid] = d.y;
            sd[q][2][tid] = d.z;
        }
    }
    __syncthreads();
    if (!active) return;
    const int t0 = cstart[c] - tb, t1 = cstart[c + 1] - tb;
    for (int e = t0; e < t1; e++) {
        const uint32_t loc = stl[e];
        float bb[12], dd[12], w[4];
        int slot[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            slot[j] = (loc >> (4 * j)) & 15u;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                bb[3 * j + k] = sb[slot[j]][k][tid];
                dd[3 * j + k] = sd[slot[j]][k][tid];
            }
            w[j] = sb[slot[j]][3][tid];
        }
        float q[9];
#pragma unroll
        for (int k = 0; k < 9; k++) q[k] = sq[k][e];
        float l[2] = { 0.0f, 0.0f };
        if (!FIRST) {
            l[0] = sl[0][e];
            l[1] = sl[1][e];
        }
        float dx[12];
        project_tet_split<float>(bb, dd, w, q, sq[9][e], sq[10][e], sq[11][e], l, dx);
        ta.lam[tb + e] = make_float2(l[0], l[1]);
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int k = 0; k < 3; k++) sd[slot[j]][k][tid] = dd[3 * j + k] + dx[3 * j + k];
    }
#pragma unroll
    for (int q = 0; q < 8; q++) {
        if (vid[q] < 0) continue;
        pos[2 * (size_t)vid[q] + 1] = make_float4(sd[q][0][tid], sd[q][1][tid], sd[q][2][tid], 0.f);
    }
}

// ---------------------------------------------------------------------------
// Quad cluster Gauss-Seidel: the cluster order of k_gs_cluster, with four lanes
// (one DPP quad) per cluster.  For each element, lane j owns the element's
// vertex j: it forms its row of D = [x0-x3, x1-x3, x2-x3] and its share of
// F = I + D Q, the quad sums F, and every lane evaluates C_D, C_H and the 2x2
// system redundantly; lane j then computes the gradient column of its own
// vertex (vertex 3 through Q's fourth column -(Q_0 + Q_1 + Q_2)), the quad sums
// the three entries of the system matrix, and lane j corrects its own vertex.
// A quad sum (x + swap1) + swap2 is the same sum in all four lanes, so the
// lanes agree bit for bit.  The shorter per-lane chain and four times the
// lanes per colour are the point: a colour holds only ~1/8 of the clusters.
// ---------------------------------------------------------------------------

__device__ __forceinline__ float dpp_xor1(float x)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0xB1, 0xF, 0xF, true));
}
__device__ __forceinline__ float dpp_xor2(float x)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x4E, 0xF, 0xF, true));
}
// Sum of a * b over the four lanes of a quad, (x + swap1) + swap2, the same
// bits in all four lanes.  The product is rounded in every lane first:
// contracted into an FMA, each lane would add its own unrounded product to its
// partners' rounded ones, and the lanes would no longer hold the same sum.
__device__ __forceinline__ float quad_sum_of_product(float a, float b)
{
#pragma clang fp contract(off)
    const float x = a * b;
    const float y = x + dpp_xor1(x);
    return y + dpp_xor2(y);
}
__device__ __forceinline__ float quad_lane3(float x)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0xFF, 0xF, 0xF, true));
}

// BLOCK threads per workgroup = BLOCK / 4 clusters; the launch shape is a tuning choice.
// MT bounds the elements of a cluster: 8 in general, 6 where every cluster is
// one hexahedral cell, which leaves a wave's element planes a quarter smaller
// and fits more waves in a CU's LDS.
// LEAN keeps only what the solve's dependent chain reads in LDS: vertex 3's
// column of Q is then summed in the quad, and the vertex bases are fetched
// from the staging lanes' registers before the chain (more instructions, less
// LDS, more waves per CU).
template <int BLOCK, int MT, bool LEAN>
struct QuadLds {
    static constexpr int kClusters = BLOCK / 4;
    static constexpr int kMaxTets = MT;           // per cluster
    static constexpr int kStage = kClusters * kMaxTets;
    static constexpr int kCol = kClusters + 1;    // padded column stride (LDS banks)
    static constexpr int kQRows = LEAN ? 9 : 12;
    float sb[LEAN ? 1 : 8 * 4 * kCol];    // [slot][xyz, w][cluster]
    float sd[8 * 3 * kCol];               // [slot][xyz][cluster]
    // row 3 j + c: Q(j, c), the gradient weights of element vertex j (vertex 3:
    // -(Q(0, c) + Q(1, c) + Q(2, c)), rows 9..11, not kept when LEAN); then
    // alpha tilde (2 rows) and gamma
    float sq[kQRows + 3][kStage];
    float sl[2][kStage];       // lambda
    uint32_t stl[kStage];      // local vertex slots
};

// The unrolled step loop covers every element of a cluster.
static_assert(kClusterTets == 8 && kClusterVerts == 8, "cluster shape of build_clusters");

// Where the in-step motion (delta) of vertex v lies: after its base in the
// interleaved vertex record, or, while a Gauss-Seidel iteration loop runs in
// the split layout, in its own array (dsep), where the deltas of neighbouring
// vertices share 32-byte write sectors.
template <bool SPLIT>
__device__ __forceinline__ float4* delta_at(float4* pos, float4* dsep, int v)
{
    return SPLIT ? dsep + v : pos + 2 * (size_t)v + 1;
}

// One chunk of BLOCK / 4 consecutive clusters starting at cluster cb (clusters
// below cend), projected by one workgroup and written back.  Each wave owns 16
// consecutive clusters and everything it stages in LDS (their vertex columns
// and, in its own slice of the element planes, their elements), so the waves
// need no workgroup barrier: a wave's LDS operations complete in order.
template <bool FIRST, int BLOCK, int MT, bool LEAN, bool SPLIT>
__device__ __forceinline__ void quad_chunk(QuadLds<BLOCK, MT, LEAN>& L, const TetArrays& ta, const int4* __restrict__ cverts,
                                           const int* __restrict__ cstart, const uint32_t* __restrict__ tloc,
                                           float4* __restrict__ pos, float4* __restrict__ dsep,
                                           const float4* __restrict__ dbase, int cb, int cend)
{
    constexpr int kQuadBlock = BLOCK;
    using Lds = QuadLds<BLOCK, MT, LEAN>;
    constexpr int kQuadStage = Lds::kStage;
    constexpr int kQuadCol = Lds::kCol;
    constexpr int kMaxTets = Lds::kMaxTets;
    constexpr int kQR = Lds::kQRows;   // first row after Q's
    float* sb = L.sb;
    float* sd = L.sd;
    auto& sq = L.sq;
    auto& sl = L.sl;
    uint32_t* stl = L.stl;
    constexpr int kWaveClusters = 16;                      // 64 lanes, 4 per cluster
    constexpr int kWaveStage = kWaveClusters * kMaxTets;    // element slots of one wave
    constexpr int kPer = (kWaveStage + 63) / 64;           // elements staged per lane
    static_assert(kQuadStage == (kQuadBlock / 64) * kWaveStage, "element planes split into wave slices");
    const int tid = threadIdx.x;
    const int cl = tid >> 2, j = tid & 3, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int c = cb + cl;
    const bool active = c < cend;
    // clusters of this wave below cend (the same in every lane)
    const int nwc = min(kWaveClusters, cend - (cb + wave * kWaveClusters));
    if (nwc <= 0) return;  // no workgroup barrier follows
    // The cluster table and ranges first; the wave's element range then comes
    // from them (its clusters are consecutive), and its element data and the
    // vertex records are loaded together.
    int4 va = make_int4(-1, -1, -1, -1), vb = va;
    int cs0 = 0, cs1 = 0;
    if (active) {
        va = cverts[2 * (size_t)c];
        vb = cverts[2 * (size_t)c + 1];
        cs0 = cstart[c];
        cs1 = cstart[c + 1];
    }
    const int tb = __builtin_amdgcn_readfirstlane(cs0);              // lane 0's cluster is the wave's first
    const int n = __builtin_amdgcn_readlane(cs1, 4 * nwc - 1) - tb;  // end of its last active cluster
    const int sbase = wave * kWaveStage;                             // the wave's slice of the element planes
    float eq[kPer][12], el[kPer][2];
    uint32_t etl[kPer];
#pragma unroll
    for (int r = 0; r < kPer; r++) {
        const int i = lane + r * 64;
        if (i < n) {
#pragma unroll
            for (int k = 0; k < 9; k++) eq[r][k] = ta.Q[(size_t)k * ta.qstride + tb + i];
            const float2 a = ta.at[tb + i];
            eq[r][9] = a.x;
            eq[r][10] = a.y;
            eq[r][11] = ta.gamma[tb + i];
            if (!FIRST) {
                const float2 lv = ta.lam[tb + i];
                el[r][0] = lv.x;
                el[r][1] = lv.y;
            }
            etl[r] = tloc[tb + i];
        }
    }
    // this lane stages and writes back slots j and j + 4
    const int v_lo = j == 0 ? va.x : j == 1 ? va.y : j == 2 ? va.z : va.w;
    const int v_hi = j == 0 ? vb.x : j == 1 ? vb.y : j == 2 ? vb.z : vb.w;
    float4 blo = make_float4(0.f, 0.f, 0.f, 0.f), dlo = blo, bhi = blo, dhi = blo;
    // (the interleaved delta is written as an offset from its base's address,
    // which the loads fold into one address)
    if (v_lo >= 0) {
        if constexpr (SPLIT) {
            blo = dbase[v_lo];
            dlo = dsep[v_lo];
        } else {
            blo = pos[2 * (size_t)v_lo];
            dlo = pos[2 * (size_t)v_lo + 1];
        }
    }
    if (v_hi >= 0) {
        if constexpr (SPLIT) {
            bhi = dbase[v_hi];
            dhi = dsep[v_hi];
        } else {
            bhi = pos[2 * (size_t)v_hi];
            dhi = pos[2 * (size_t)v_hi + 1];
        }
    }
#pragma unroll
    for (int r = 0; r < kPer; r++) {
        const int i = sbase + lane + r * 64;
        if (lane + r * 64 < n) {
            // Q is column-major (eq[3 c + r] = Q(r, c)); row 3 j + c of sq holds Q(j, c)
#pragma unroll
            for (int cc = 0; cc < 3; cc++) {
                const float q0 = eq[r][3 * cc], q1 = eq[r][3 * cc + 1], q2 = eq[r][3 * cc + 2];
                sq[cc][i] = q0;
                sq[3 + cc][i] = q1;
                sq[6 + cc][i] = q2;
                if constexpr (!LEAN) sq[9 + cc][i] = -(q0 + q1 + q2);
            }
            sq[kQR][i] = eq[r][9];
            sq[kQR + 1][i] = eq[r][10];
            sq[kQR + 2][i] = eq[r][11];
            if (!FIRST) {
                sl[0][i] = el[r][0];
                sl[1][i] = el[r][1];
            }
            stl[i] = etl[r];
        }
    }
    if (v_lo >= 0) {
        if constexpr (!LEAN) {
            sb[(j * 4 + 0) * kQuadCol + cl] = blo.x;
            sb[(j * 4 + 1) * kQuadCol + cl] = blo.y;
            sb[(j * 4 + 2) * kQuadCol + cl] = blo.z;
            sb[(j * 4 + 3) * kQuadCol + cl] = blo.w;
        }
        sd[(j * 3 + 0) * kQuadCol + cl] = dlo.x;
        sd[(j * 3 + 1) * kQuadCol + cl] = dlo.y;
        sd[(j * 3 + 2) * kQuadCol + cl] = dlo.z;
    }
    if (v_hi >= 0) {
        const int q = j + 4;
        if constexpr (!LEAN) {
            sb[(q * 4 + 0) * kQuadCol + cl] = bhi.x;
            sb[(q * 4 + 1) * kQuadCol + cl] = bhi.y;
            sb[(q * 4 + 2) * kQuadCol + cl] = bhi.z;
            sb[(q * 4 + 3) * kQuadCol + cl] = bhi.w;
        }
        sd[(q * 3 + 0) * kQuadCol + cl] = dhi.x;
        sd[(q * 3 + 1) * kQuadCol + cl] = dhi.y;
        sd[(q * 3 + 2) * kQuadCol + cl] = dhi.z;
    }
    // the wave reads only what it staged (LDS operations of one wave complete in order)
    __builtin_amdgcn_wave_barrier();
    // A quad is active or inactive as a whole (an inactive quad has no
    // elements), so DPP reads stay inside active quads.
    const int t0 = active ? sbase + cs0 - tb : 0;
    const int nte = active ? cs1 - cs0 : 0;   // at most kMaxTets
    // What stays fixed during the solve is read once per element, before the
    // dependent chain: this lane's slot (as its LDS offset in sd), the base part
    // of its row of D = (b_j - b_3) + (d_j - d_3), and its inverse mass.  LEAN:
    // slot s of a cluster is staged by lane s & 3 of its quad (low or high
    // record) and is fetched from that lane's registers.
    int soff[kMaxTets];
    float Db[kMaxTets][3], wv[kMaxTets];
    const float blv[4] = { blo.x, blo.y, blo.z, blo.w }, bhv[4] = { bhi.x, bhi.y, bhi.z, bhi.w };
#pragma unroll
    for (int k = 0; k < kMaxTets; k++) {
        soff[k] = cl;
        Db[k][0] = Db[k][1] = Db[k][2] = 0.f;
        wv[k] = 0.f;
        if (k < nte) {
            const int slot = (int)((stl[t0 + k] >> (4 * j)) & 15u);
            soff[k] = slot * 3 * kQuadCol + cl;
            float bb[4];
            if constexpr (LEAN) {
                const int src = ((lane & ~3) | (slot & 3)) << 2;   // byte address of the owning lane
#pragma unroll
                for (int kk = 0; kk < 4; kk++) {
                    const float lo = __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(blv[kk])));
                    const float hi = __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(bhv[kk])));
                    bb[kk] = slot < 4 ? lo : hi;
                }
            } else {
#pragma unroll
                for (int kk = 0; kk < 4; kk++) bb[kk] = sb[(slot * 4 + kk) * kQuadCol + cl];
            }
#pragma unroll
            for (int kk = 0; kk < 3; kk++) Db[k][kk] = bb[kk] - quad_lane3(bb[kk]);
            wv[k] = bb[3];
        }
    }
#pragma unroll
    for (int k = 0; k < kMaxTets; k++) {
        if (k < nte) {
            const int e = t0 + k;
            const int so = soff[k];
            const float w = wv[k];
            // every LDS read of the step is independent of the others: one round trip
            float dd[3], qj[3];
#pragma unroll
            for (int kk = 0; kk < 3; kk++) dd[kk] = sd[so + kk * kQuadCol];
#pragma unroll
            for (int cc = 0; cc < 3; cc++) {
                if constexpr (LEAN) {
                    // vertex 3: -((Q(0, c) + Q(1, c)) + Q(2, c)), the quad sum with lane 3 adding zero
                    const float x = j < 3 ? sq[3 * j + cc][e] : 0.f;
                    const float y = x + dpp_xor1(x);
                    const float sum = y + dpp_xor2(y);
                    qj[cc] = j < 3 ? x : -sum;
                } else {
                    qj[cc] = sq[3 * j + cc][e];
                }
            }
            const float gamma = sq[kQR + 2][e];
            const float atD = sq[kQR][e], atH = sq[kQR + 1][e];
            const float l0 = FIRST ? 0.f : sl[0][e], l1 = FIRST ? 0.f : sl[1][e];
            // this lane's row of D (zero for vertex 3)
            float D[3];
#pragma unroll
            for (int kk = 0; kk < 3; kk++) D[kk] = Db[k][kk] + (dd[kk] - quad_lane3(dd[kk]));
            float F[9];
#pragma unroll
            for (int cc = 0; cc < 3; cc++)
#pragma unroll
                for (int r = 0; r < 3; r++) {
                    const float sum = quad_sum_of_product(D[r], qj[cc]);
                    F[3 * cc + r] = r == cc ? sum + 1.f : sum;
                }
            // C_D, C_H as element_constraints, with both branches of C_H evaluated
            float ss = 0.f;
#pragma unroll
            for (int i = 0; i < 9; i++) ss += F[i] * F[i];
            const float CD = xsqrt<float>(ss);
            float Fc[9];
            Fc[0] = F[4] * F[8] - F[5] * F[7];
            Fc[1] = F[5] * F[6] - F[3] * F[8];
            Fc[2] = F[3] * F[7] - F[4] * F[6];
            Fc[3] = F[7] * F[2] - F[8] * F[1];
            Fc[4] = F[8] * F[0] - F[6] * F[2];
            Fc[5] = F[6] * F[1] - F[7] * F[0];
            Fc[6] = F[1] * F[5] - F[2] * F[4];
            Fc[7] = F[2] * F[3] - F[0] * F[5];
            Fc[8] = F[0] * F[4] - F[1] * F[3];
            // J = det F by cofactors along column 0 (Fc[0..2] are its cofactors)
            const float J = F[0] * Fc[0] + F[1] * Fc[1] + F[2] * Fc[2];
            const float a = J - 1.f;
            const float ch_log = -gamma + xlog<float>(J);
            const float ch_series = -gamma + a - a * a * 0.5f + a * a * a * (1.f / 3.f);
            const float CH = J >= 1.f ? ch_log : ch_series;
            const float inv_C = xrcp<float>(CD);
            const float fac = (J >= 1.f) ? xrcp<float>(J) : 1.f - a + a * a;
            float gD[3], gH[3];
#pragma unroll
            for (int r = 0; r < 3; r++) {
                gD[r] = inv_C * (F[r] * qj[0] + F[3 + r] * qj[1] + F[6 + r] * qj[2]);
                gH[r] = fac * (Fc[r] * qj[0] + Fc[3 + r] * qj[1] + Fc[6 + r] * qj[2]);
            }
            const float s00 = quad_sum_of_product(w, gD[0] * gD[0] + gD[1] * gD[1] + gD[2] * gD[2]);
            const float s01 = quad_sum_of_product(w, gD[0] * gH[0] + gD[1] * gH[1] + gD[2] * gH[2]);
            const float s11 = quad_sum_of_product(w, gH[0] * gH[0] + gH[1] * gH[1] + gH[2] * gH[2]);
            const float A00 = atD + s00, A11 = atH + s11, A01 = s01;
            const float b0 = -CD - atD * l0;
            const float b1 = -CH - atH * l1;
            // SPD in exact arithmetic; a non-positive FP32 determinant (slivers) skips
            // the projection, as in project_tet_F
            const float det = A00 * A11 - A01 * A01;
            const float inv_det = det > 0.f ? xrcp<float>(det) : 0.f;
            const float d0 = (b0 * A11 - b1 * A01) * inv_det;
            const float d1 = (b1 * A00 - b0 * A01) * inv_det;
            // the four lanes hold the same multipliers and store the same value
            sl[0][e] = l0 + d0;
            sl[1][e] = l1 + d1;
#pragma unroll
            for (int kk = 0; kk < 3; kk++) sd[so + kk * kQuadCol] = dd[kk] + w * (gD[kk] * d0 + gH[kk] * d1);
            // the next element's reads of other lanes' slots follow these writes in
            // program order (LDS operations of one wave complete in order)
            __builtin_amdgcn_wave_barrier();
        }
    }
    if (v_lo >= 0) {
        const float4 d = make_float4(sd[(j * 3 + 0) * kQuadCol + cl], sd[(j * 3 + 1) * kQuadCol + cl],
                                     sd[(j * 3 + 2) * kQuadCol + cl], 0.f);
        if constexpr (SPLIT) dsep[v_lo] = d;
        else pos[2 * (size_t)v_lo + 1] = d;
    }
    if (v_hi >= 0) {
        const int qs = j + 4;
        const float4 d = make_float4(sd[(qs * 3 + 0) * kQuadCol + cl], sd[(qs * 3 + 1) * kQuadCol + cl],
                                     sd[(qs * 3 + 2) * kQuadCol + cl], 0.f);
        if constexpr (SPLIT) dsep[v_hi] = d;
        else pos[2 * (size_t)v_hi + 1] = d;
    }
    // multipliers of the wave's tet range, coalesced
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int r = 0; r < kPer; r++) {
        const int i = lane + r * 64;
        if (i < n) ta.lam[tb + i] = make_float2(sl[0][sbase + i], sl[1][sbase + i]);
    }
}

template <bool FIRST, int BLOCK, int MT, bool LEAN, bool SPLIT>
__global__ __launch_bounds__(BLOCK) void k_gs_cluster_quad(TetArrays ta, const int4* __restrict__ cverts,
                                                          const int* __restrict__ cstart,
                                                          const uint32_t* __restrict__ tloc,
                                                          float4* __restrict__ pos, float4* __restrict__ dsep,
                                                          const float4* __restrict__ dbase, int cbeg, int cend,
                                                          int xcd)
{
    __shared__ QuadLds<BLOCK, MT, LEAN> L;
    const int cb = cbeg + xcd_item(blockIdx.x, gridDim.x, xcd != 0) * QuadLds<BLOCK, MT, LEAN>::kClusters;
    quad_chunk<FIRST, BLOCK, MT, LEAN, SPLIT>(L, ta, cverts, cstart, tloc, pos, dsep, dbase, cb, cend);
}

// The split layout for a Gauss-Seidel iteration loop: the deltas copied out of
// the vertex records before it (dsep[0, nv)) and back after it, and the bases,
// which the loop only reads, copied once into dsep[nv, 2 nv).
__global__ __launch_bounds__(kBlock) void k_split_deltas(const float4* __restrict__ pos, float4* __restrict__ dsep,
                                                         int nv)
{
    const int v = blockIdx.x * kBlock + threadIdx.x;
    if (v < nv) {
        dsep[v] = pos[2 * (size_t)v + 1];
        dsep[(size_t)nv + v] = pos[2 * (size_t)v];
    }
}
__global__ __launch_bounds__(kBlock) void k_join_deltas(float4* __restrict__ pos, const float4* __restrict__ dsep,
                                                        int nv)
{
    const int v = blockIdx.x * kBlock + threadIdx.x;
    if (v < nv) pos[2 * (size_t)v + 1] = dsep[v];
}

// ---------------------------------------------------------------------------
// Jacobi: every projection reads the same positions; corrections are summed
// into acc (4 floats per vertex) with hardware float atomics, then applied.
// ---------------------------------------------------------------------------
template <bool FIRST>
__global__ __launch_bounds__(kBlock) void k_jacobi_project(TetArrays ta, const float4* __restrict__ pos,
                                                          float* __restrict__ acc, int nt)
{
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= nt) return;
    const int4 t = ta.tet[e];
    Gather g;
    gather4(pos, t, g);
    float q[9];
    float2 at;
    float gm;
    load_tet(ta, e, q, at, gm);
    float l[2] = { 0.0f, 0.0f };
    if (!FIRST) {
        const float2 lv = ta.lam[e];
        l[0] = lv.x;
        l[1] = lv.y;
    }
    float dx[12];
    project_tet_split<float>(g.b, g.d, g.w, q, at.x, at.y, gm, l, dx);
    ta.lam[e] = make_float2(l[0], l[1]);
    const int vid[4] = { t.x, t.y, t.z, t.w };
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) unsafeAtomicAdd(acc + 4 * (size_t)vid[i] + k, dx[3 * i + k]);
}

__global__ __launch_bounds__(kBlock) void k_jacobi_apply(float4* __restrict__ pos, float4* __restrict__ acc, int nv)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= nv) return;
    float4 d = pos[2 * (size_t)i + 1];
    const float4 a = acc[i];
    d.x += a.x;
    d.y += a.y;
    d.z += a.z;
    pos[2 * (size_t)i + 1] = d;
    acc[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// ---------------------------------------------------------------------------
// Patch Jacobi.  A workgroup owns a patch of kPatch consecutive (Morton-sorted)
// tets: it stages the patch's distinct vertices in LDS, each thread projects
// one tet reading LDS, the corrections are summed per local vertex in LDS, and
// then every local vertex is written once: interior vertices (no tet outside
// the patch touches them, so no other workgroup reads them this iteration)
// are updated in place, shared vertices leave their partial sum in a fixed
// slot that k_jacobi_shared adds in fixed slot order.
//   DET   sum corrections with a fixed-order gather over the patch's local
//         incidence lists instead of LDS float atomics (bitwise reproducible)
//   CON   fold in the contact corrections accumulated in cacc
// ---------------------------------------------------------------------------
constexpr int kPatch = 256;

// Sum the corrections of one local vertex over its incidence entries s0..s1-1
// (entry = 4 * local tet + corner; plane k of red holds component k of every
// tet's corrections) in entry order.  Four entries at a time: their indices,
// then their values, are read together, and added in the same order.
__device__ __forceinline__ void gather_incidence(const uint16_t* il, const float* red, int s0, int s1, float& a0,
                                                 float& a1, float& a2)
{
    int s = s0;
    for (; s + 4 <= s1; s += 4) {
        int off[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int ent = il[s + u];
            off[u] = (3 * (ent & 3)) * kPatch + (ent >> 2);
        }
        float r[4][3];
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
            for (int k = 0; k < 3; k++) r[u][k] = red[off[u] + k * kPatch];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            a0 += r[u][0];
            a1 += r[u][1];
            a2 += r[u][2];
        }
    }
    for (; s < s1; s++) {
        const int ent = il[s];
        const int off = (3 * (ent & 3)) * kPatch + (ent >> 2);
        a0 += red[off];
        a1 += red[off + kPatch];
        a2 += red[off + 2 * kPatch];
    }
}

template <bool FIRST, bool DET, bool CON>
__global__ __launch_bounds__(kPatch) void k_jacobi_patch(TetArrays ta, const uint2* __restrict__ tloc,
                                                        const int* __restrict__ pvoff, const int* __restrict__ pverts,
                                                        const int* __restrict__ pslot, const int* __restrict__ inc_off,
                                                        const uint16_t* __restrict__ inc, float4* __restrict__ pos,
                                                        float4* __restrict__ partial, float4* __restrict__ cacc,
                                                        int nt, int maxl, int xcd)
{
    extern __shared__ float4 smem[];
    float4* lb = smem;                                   // base, inverse mass
    float* ld = reinterpret_cast<float*>(smem + maxl);   // in-step motion, 3 per local vertex
    float* red = ld + 3 * maxl;                          // DET: 12 planes of kPatch; else 3 per local vertex
    uint32_t* linc = reinterpret_cast<uint32_t*>(red + 12 * kPatch);  // DET: the patch's incidence list
    const int tid = threadIdx.x;
    const int p = xcd_item(blockIdx.x, gridDim.x, xcd != 0);
    const int v0 = pvoff[p], nl = pvoff[p + 1] - v0;
    // element data first, so its latency overlaps the vertex staging
    const int e = p * kPatch + tid;
    const bool has = e < nt;
    uint2 lc = make_uint2(0u, 0u);
    float q[9];
    float2 at = make_float2(0.f, 0.f);
    float gm = 0.f;
    float l[2] = { 0.0f, 0.0f };
    if (has) {
        lc = tloc[e];
        load_tet(ta, e, q, at, gm);
        if (!FIRST) {
            const float2 lv = ta.lam[e];
            l[0] = lv.x;
            l[1] = lv.y;
        }
    }
    // DET: the incidence list (4 entries per tet, 16 bits each) goes to LDS with
    // the other loads, so the gather below walks LDS only
    if (DET) {
        // 4-byte LDS stores: the list's LDS offset is a multiple of 4, not always of 8
        const uint2 e = reinterpret_cast<const uint2*>(inc + (size_t)p * 4 * kPatch)[tid];
        linc[2 * tid] = e.x;
        linc[2 * tid + 1] = e.y;
    }
    // this thread's first local vertex: its id, slot and incidence range are
    // loaded with the staging, so the write-back phase starts without a global
    // round trip
    const int ioff = v0 + p;  // this patch's nl+1 incidence offsets
    int g_first = 0, slot_first = 0, s0_first = 0, s1_first = 0;
    if (tid < nl) {
        slot_first = pslot[v0 + tid];
        if (DET) {
            s0_first = inc_off[ioff + tid];
            s1_first = inc_off[ioff + tid + 1];
        }
    }
    for (int i = tid; i < nl; i += kPatch) {
        const int g = pverts[v0 + i];
        if (i == tid) g_first = g;
        lb[i] = pos[2 * (size_t)g];
        const float4 d = pos[2 * (size_t)g + 1];
        ld[3 * i] = d.x;
        ld[3 * i + 1] = d.y;
        ld[3 * i + 2] = d.z;
        if (!DET) red[3 * i] = red[3 * i + 1] = red[3 * i + 2] = 0.f;
    }
    __syncthreads();
    if (has) {
        const int sl[4] = { (int)(lc.x & 0xffffu), (int)(lc.x >> 16), (int)(lc.y & 0xffffu), (int)(lc.y >> 16) };
        float bb[12], dd[12], w[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const float4 b = lb[sl[j]];
            bb[3 * j] = b.x;
            bb[3 * j + 1] = b.y;
            bb[3 * j + 2] = b.z;
            w[j] = b.w;
#pragma unroll
            for (int k = 0; k < 3; k++) dd[3 * j + k] = ld[3 * sl[j] + k];
        }
        float dx[12];
        project_tet_split<float>(bb, dd, w, q, at.x, at.y, gm, l, dx);
        ta.lam[e] = make_float2(l[0], l[1]);
        if (DET) {
#pragma unroll
            for (int k = 0; k < 12; k++) red[k * kPatch + tid] = dx[k];
        } else {
#pragma unroll
            for (int j = 0; j < 4; j++)
#pragma unroll
                for (int k = 0; k < 3; k++) atomicAdd(red + 3 * sl[j] + k, dx[3 * j + k]);
        }
    }
    __syncthreads();
    int g = g_first, slot = slot_first, s0 = s0_first, s1 = s1_first;
    for (int i = tid; i < nl; i += kPatch) {
        if (i != tid) {
            g = pverts[v0 + i];
            slot = pslot[v0 + i];
            if (DET) {
                s0 = inc_off[ioff + i];
                s1 = inc_off[ioff + i + 1];
            }
        }
        float a0, a1, a2;
        if (DET) {
            a0 = a1 = a2 = 0.f;
            gather_incidence(reinterpret_cast<const uint16_t*>(linc), red, s0, s1, a0, a1, a2);
        } else {
            a0 = red[3 * i];
            a1 = red[3 * i + 1];
            a2 = red[3 * i + 2];
        }
        if (slot < 0) {
            if (CON) {
                const float4 c = cacc[g];
                a0 += c.x;
                a1 += c.y;
                a2 += c.z;
                cacc[g] = make_float4(0.f, 0.f, 0.f, 0.f);
            }
            pos[2 * (size_t)g + 1] = make_float4(ld[3 * i] + a0, ld[3 * i + 1] + a1, ld[3 * i + 2] + a2, 0.f);
        } else {
            partial[slot] = make_float4(a0, a1, a2, 0.f);
        }
    }
}

template <bool CON>
__global__ __launch_bounds__(kBlock) void k_jacobi_shared(const int* __restrict__ shared, const int* __restrict__ soff,
                                                         const float4* __restrict__ partial,
                                                         float4* __restrict__ cacc, float4* __restrict__ pos, int ns)
{
    const int j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= ns) return;
    const int g = shared[j];
    float a0 = 0.f, a1 = 0.f, a2 = 0.f;
    // the vertex's slots are contiguous, in patch order (fixed summation order);
    // four loads are issued together
    const int k0 = soff[j], k1 = soff[j + 1];
    int k = k0;
    for (; k + 4 <= k1; k += 4) {
        float4 v[4];
#pragma unroll
        for (int u = 0; u < 4; u++) v[u] = partial[k + u];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            a0 += v[u].x;
            a1 += v[u].y;
            a2 += v[u].z;
        }
    }
    for (; k < k1; k++) {
        const float4 v = partial[k];
        a0 += v.x;
        a1 += v.y;
        a2 += v.z;
    }
    if (CON) {
        const float4 c = cacc[g];
        a0 += c.x;
        a1 += c.y;
        a2 += c.z;
        cacc[g] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    float4 d = pos[2 * (size_t)g + 1];
    d.x += a0;
    d.y += a1;
    d.z += a2;
    pos[2 * (size_t)g + 1] = d;
}

__global__ __launch_bounds__(kBlock) void k_fill_zero(uint32_t* __restrict__ p, size_t n)
{
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// ---------------------------------------------------------------------------
// FP64 code below (residual, energy, signed distances, detection, contact and
// attachment projection) is compiled without FMA contraction: every product is
// rounded before it is added, as in the FP64 restatement of the reference
// arithmetic, so that detection from the same state takes the same branches
// (Frank-Wolfe's vertex choice compares dot products that can tie exactly).
#pragma clang fp contract(off)
// ---------------------------------------------------------------------------
// FP64 residual / energy (XPBDSolver.hpp:327-391, XPBDMeshObject.cpp:537-549).
// The FP32 state is promoted and the constraint maths evaluated in double.
// Partial sums use a fixed tree, so results are reproducible for a fixed grid.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void disp_f64(const float4* __restrict__ pos, const float4* __restrict__ lo, int v,
                                         bool with_delta, double* u)
{
    const float4 b = pos[2 * (size_t)v];
    const float4 l = lo[v];
    u[0] = (double)b.x + (double)l.x;
    u[1] = (double)b.y + (double)l.y;
    u[2] = (double)b.z + (double)l.z;
    if (with_delta) {
        const float4 d = pos[2 * (size_t)v + 1];
        u[0] += (double)d.x;
        u[1] += (double)d.y;
        u[2] += (double)d.z;
    }
}

__device__ __forceinline__ void eval_tet_f64(const TetArrays& ta, const float4* pos, const float4* lo, bool with_delta,
                                             int e, double* CD, double* CH)
{
    const int4 t = ta.tet[e];
    double u0[3], u1[3], u2[3], u3[3];
    disp_f64(pos, lo, t.x, with_delta, u0);
    disp_f64(pos, lo, t.y, with_delta, u1);
    disp_f64(pos, lo, t.z, with_delta, u2);
    disp_f64(pos, lo, t.w, with_delta, u3);
    double q[9];
#pragma unroll
    for (int k = 0; k < 9; k++) q[k] = ta.Q[(size_t)k * ta.qstride + e];
    double F[9];
    deformation_gradient<double>(u0, u1, u2, u3, q, F);
    double J;
    element_constraints<double>(F, (double)ta.gamma[e], CD, CH, &J);
}

// Multiplier start of the learned step: every element's multipliers at the
// values that satisfy its two constraints at the current positions,
// lambda = -C / alpha~ (C_D, C_H as in DeviatoricConstraint.hpp:112-152 and
// HydrostaticConstraint.hpp:123-153, alpha~ = alpha / dt^2), with C evaluated
// in FP64 as the residual is, so that the start's residual is the rounding of
// the FP32 multipliers.  A solve started from them without the reset of
// XPBDSolver.hpp:66-77 begins at zero elastic residual.
__global__ __launch_bounds__(kBlock) void k_lambda_start(TetArrays ta, const float4* __restrict__ pos,
                                                        const float4* __restrict__ lo, int nt)
{
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= nt) return;
    double CD, CH;
    eval_tet_f64(ta, pos, lo, true, e, &CD, &CH);
    const float2 at = ta.at[e];
    ta.lam[e] = make_float2((float)(-CD / (double)at.x), (float)(-CH / (double)at.y));
}

__device__ __forceinline__ double block_sum(double v, double* sh)
{
    sh[threadIdx.x] = v;
    __syncthreads();
#pragma unroll
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    return sh[0];
}

__global__ __launch_bounds__(kBlock) void k_residual(TetArrays ta, const float4* __restrict__ pos,
                                                    const float4* __restrict__ lo, int with_delta, int nt,
                                                    double* __restrict__ part, double* __restrict__ vec)
{
    __shared__ double sh[kBlock];
    double acc = 0.0;
    for (int e = blockIdx.x * kBlock + threadIdx.x; e < nt; e += gridDim.x * kBlock) {
        double CD, CH;
        eval_tet_f64(ta, pos, lo, with_delta != 0, e, &CD, &CH);
        const float2 at = ta.at[e];
        const float2 l = ta.lam[e];
        const double r1 = CD + (double)at.x * (double)l.x;
        const double r2 = CH + (double)at.y * (double)l.y;
        if (vec) {
            vec[2 * (size_t)e] = r1;
            vec[2 * (size_t)e + 1] = r2;
        }
        acc += r1 * r1 + r2 * r2;
    }
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void k_energy(TetArrays ta, const float4* __restrict__ pos,
                                                  const float4* __restrict__ lo, int nt, double dt2,
                                                  double* __restrict__ part)
{
    __shared__ double sh[kBlock];
    double acc = 0.0;
    for (int e = blockIdx.x * kBlock + threadIdx.x; e < nt; e += gridDim.x * kBlock) {
        double CD, CH;
        eval_tet_f64(ta, pos, lo, false, e, &CD, &CH);
        const float2 at = ta.at[e];
        acc += CD * CD / ((double)at.x * dt2) + CH * CH / ((double)at.y * dt2);
    }
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void k_sum_partials(const double* __restrict__ part, int n,
                                                        double* __restrict__ out)
{
    __shared__ double sh[kBlock];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += kBlock) acc += part[i];
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) out[0] = s;
}

// ---------------------------------------------------------------------------
// Primary residual: the momentum balance of Equation (8) of the XPBD paper,
//   R_v = (x_v - xt_v) / w_v - sum grad_v C^T lambda,
// which the reference declares (XPBDSolver.hpp:127-130) and leaves empty
// (XPBDSolver.hpp:322).  The sum runs over both constraints of every element
// around v and every active contact on a face around v (gradient barycentric *
// normal, StaticDeformableCollisionConstraint.cpp:20-55).  x - xt is the
// in-step motion minus the motion the inertial prediction set, so it is formed
// exactly.  Pinned vertices read zero (their force is the support's), and so
// do attached ones: a zero-compliance attachment holds its vertex wherever the
// balance needs.  One thread per element writes grad C^T lambda at each of its
// four vertices into that vertex's slot of a vertex-ordered list; one thread
// per vertex then sums its contiguous slots and its faces' terms in a fixed
// order, so the result is reproducible.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_keep_prediction(const float4* __restrict__ pos, float4* __restrict__ pred,
                                                           int nv)
{
    const int v = blockIdx.x * kBlock + threadIdx.x;
    if (v < nv) pred[v] = pos[2 * (size_t)v + 1];
}

__global__ __launch_bounds__(kBlock) void k_element_forces(TetArrays ta, PrimaryArgs A)
{
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= A.nt) return;
    const int4 t = ta.tet[e];
    double u0[3], u1[3], u2[3], u3[3];
    disp_f64(A.pos, A.lo, t.x, true, u0);
    disp_f64(A.pos, A.lo, t.y, true, u1);
    disp_f64(A.pos, A.lo, t.z, true, u2);
    disp_f64(A.pos, A.lo, t.w, true, u3);
    double q[9];
#pragma unroll
    for (int i = 0; i < 9; i++) q[i] = ta.Q[(size_t)i * ta.qstride + e];
    double F[9], ss = 0.0, gD[12], gH[12];
    deformation_gradient<double>(u0, u1, u2, u3, q, F);
#pragma unroll
    for (int i = 0; i < 9; i++) ss += F[i] * F[i];
    // the gradients need C_D = |F| and J, not C_H
    constraint_gradients<double>(F, q, sqrt(ss), det3cm<double>(F), gD, gH);
    const float2 l = ta.lam[e];
    const int4 slot = reinterpret_cast<const int4*>(A.vt_slot)[e];
    const int sl[4] = { slot.x, slot.y, slot.z, slot.w };
#pragma unroll
    for (int c = 0; c < 4; c++) {
        double* f = A.ef + 3 * (size_t)sl[c];
#pragma unroll
        for (int i = 0; i < 3; i++) f[i] = gD[3 * c + i] * (double)l.x + gH[3 * c + i] * (double)l.y;
    }
}

__global__ __launch_bounds__(kBlock) void k_primary_residual(TetArrays ta, PrimaryArgs A)
{
    const int v = blockIdx.x * kBlock + threadIdx.x;
    if (v >= A.nv) return;
    const float4 b = A.pos[2 * (size_t)v];
    double r[3] = { 0.0, 0.0, 0.0 };
    if (A.vel[v].w == 0.f && b.w > 0.f) {
        const float4 d = A.pos[2 * (size_t)v + 1];
        const float4 p = A.pred[v];
        const double w = (double)b.w;
        r[0] = ((double)d.x - (double)p.x) / w;
        r[1] = ((double)d.y - (double)p.y) / w;
        r[2] = ((double)d.z - (double)p.z) / w;
        for (int k = A.vt_off[v]; k < A.vt_off[v + 1]; k++) {
            const double* f = A.ef + 3 * (size_t)k;
#pragma unroll
            for (int i = 0; i < 3; i++) r[i] -= f[i];
        }
        if (A.vf_off) {
            for (int k = A.vf_off[v]; k < A.vf_off[v + 1]; k++) {
                const int jc = A.vf[k], j = jc / 3, c = jc - 3 * j;
                if (!A.flag[j]) continue;
                const double bc = A.cb[3 * (size_t)j + c], lam = A.clam[j];
#pragma unroll
                for (int i = 0; i < 3; i++) r[i] -= bc * A.cn[3 * (size_t)j + i] * lam;
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 3; i++) A.R[3 * (size_t)v + i] = r[i];
}

// Consistent start of the learned step: the solve begins at
//     x0 = xt + M^-1 sum_e grad C_e^T lambda0_e
// from the kept prediction, with the forces of the multipliers lambda0 that
// k_element_forces left in the vertex's slots, summed in slot order.  The
// momentum balance M (x - xt) = grad C^T lambda then holds at the start.
__global__ __launch_bounds__(kBlock) void k_consistent_shift(float4* __restrict__ pos, const float4* __restrict__ pred,
                                                            const int* __restrict__ vt_off,
                                                            const double* __restrict__ ef, int nv)
{
    const int v = blockIdx.x * kBlock + threadIdx.x;
    if (v >= nv) return;
    double f[3] = { 0.0, 0.0, 0.0 };
    for (int k = vt_off[v]; k < vt_off[v + 1]; k++) {
#pragma unroll
        for (int i = 0; i < 3; i++) f[i] += ef[3 * (size_t)k + i];
    }
    const double w = (double)pos[2 * (size_t)v].w;
    const float4 p = pred[v];
    pos[2 * (size_t)v + 1] = make_float4((float)((double)p.x + w * f[0]), (float)((double)p.y + w * f[1]),
                                         (float)((double)p.z + w * f[2]), 0.f);
}

__global__ __launch_bounds__(kBlock) void k_zero_attached(const int* __restrict__ att_v, int natt, int nv,
                                                         double* __restrict__ R)
{
    const int a = blockIdx.x * kBlock + threadIdx.x;
    if (a >= natt) return;
    const int v = att_v[a];
    if (v < 0 || v >= nv) return;
    R[3 * (size_t)v] = 0.0;
    R[3 * (size_t)v + 1] = 0.0;
    R[3 * (size_t)v + 2] = 0.0;
}

__global__ __launch_bounds__(kBlock) void k_sum_squares(const double* __restrict__ x, size_t n,
                                                       double* __restrict__ part)
{
    __shared__ double sh[kBlock];
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock)
        acc += x[i] * x[i];
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// ---------------------------------------------------------------------------
// Signed distance fields and contact detection in FP64:
// SphereSDF.cpp:16-26, BoxSDF.cpp:19-84, GeometryUtils.cpp:6-24,92-102,
// CollisionScene.cpp:171-214 (face loop) and 253-283 (Frank-Wolfe).
// ---------------------------------------------------------------------------
struct Pose {
    int kind;
    double p[3], q[4], prm[3];
};

__device__ __forceinline__ void qrot(const double* v, const double* q, double* o)
{
    const double uv = q[0] * v[0] + q[1] * v[1] + q[2] * v[2];
    const double cx = q[1] * v[2] - q[2] * v[1];
    const double cy = q[2] * v[0] - q[0] * v[2];
    const double cz = q[0] * v[1] - q[1] * v[0];
    const double s = 2 * q[3] * q[3] - 1;
    o[0] = 2 * uv * q[0] + s * v[0] + 2 * q[3] * cx;
    o[1] = 2 * uv * q[1] + s * v[1] + 2 * q[3] * cy;
    o[2] = 2 * uv * q[2] + s * v[2] + 2 * q[3] * cz;
}

__device__ __forceinline__ void to_body(const Pose& P, const double* x, double* xb)
{
    const double n2 = P.q[0] * P.q[0] + P.q[1] * P.q[1] + P.q[2] * P.q[2] + P.q[3] * P.q[3];
    const double qi[4] = { -P.q[0] / n2, -P.q[1] / n2, -P.q[2] / n2, P.q[3] / n2 };
    const double rel[3] = { x[0] - P.p[0], x[1] - P.p[1], x[2] - P.p[2] };
    qrot(rel, qi, xb);
}

__device__ double sdf_eval(const Pose& P, const double* x)
{
    if (P.kind == 1) {
        const double d0 = x[0] - P.p[0], d1 = x[1] - P.p[1], d2 = x[2] - P.p[2];
        return sqrt(d0 * d0 + d1 * d1 + d2 * d2) - P.prm[0];
    }
    double xb[3];
    to_body(P, x, xb);
    double q[3], qc[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        q[k] = fabs(xb[k]) - 0.5 * P.prm[k];
        qc[k] = q[k] > 0.0 ? q[k] : 0.0;
    }
    const double mq = fmax(q[0], fmax(q[1], q[2]));
    return sqrt(qc[0] * qc[0] + qc[1] * qc[1] + qc[2] * qc[2]) + (mq < 0.0 ? mq : 0.0);
}

__device__ void sdf_grad(const Pose& P, const double* x, double* g)
{
    if (P.kind == 1) {
        const double d[3] = { x[0] - P.p[0], x[1] - P.p[1], x[2] - P.p[2] };
        const double n = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
#pragma unroll
        for (int k = 0; k < 3; k++) g[k] = n > 0 ? d[k] / n : d[k];
        return;
    }
    double xb[3];
    to_body(P, x, xb);
    double q[3], qc[3], sg[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        q[k] = fabs(xb[k]) - P.prm[k] / 2.0;
        qc[k] = q[k] > 0.0 ? q[k] : 0.0;
        sg[k] = xb[k] > 0 ? 1.0 : -1.0;
    }
    const double mq = fmax(q[0], fmax(q[1], q[2]));
    const double md = mq < 0.0 ? mq : 0.0;
    double go[3] = { sg[0] * qc[0], sg[1] * qc[1], sg[2] * qc[2] };
    const double n = sqrt(go[0] * go[0] + go[1] * go[1] + go[2] * go[2]);
    if (n > 0) {
#pragma unroll
        for (int k = 0; k < 3; k++) go[k] /= n;
    }
    double gb[3];
#pragma unroll
    for (int k = 0; k < 3; k++) gb[k] = go[k] + sg[k] * (q[k] == md ? 1.0 : 0.0);
    qrot(gb, P.q, g);
}

// world position rest + base + low part (+ in-step motion during a solve)
__device__ __forceinline__ void world_pos(const double* rest, const float4* pos, const float4* lo, int v,
                                          bool with_delta, double* x)
{
    double u[3];
    disp_f64(pos, lo, v, with_delta, u);
    x[0] = rest[3 * (size_t)v] + u[0];
    x[1] = rest[3 * (size_t)v + 1] + u[1];
    x[2] = rest[3 * (size_t)v + 2] + u[2];
}

__global__ __launch_bounds__(kBlock) void k_detect(const int4* __restrict__ faces, int nf,
                                                  const double* __restrict__ rest,
                                                  const float4* __restrict__ pos, const float4* __restrict__ lo,
                                                  const SdfDev* __restrict__ sdf,
                                                  const unsigned long long* __restrict__ step, double dt,
                                                  int* __restrict__ flag, double* __restrict__ cp,
                                                  double* __restrict__ cn, double* __restrict__ cb)
{
    const int j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= nf) return;
    const int4 f = faces[j];
    const SdfDev S = sdf[f.w];
    int hit = 0;
    if (S.kind != 0) {
        Pose P;
        P.kind = S.kind;
        const double t = (double)step[0] * dt;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            P.p[k] = S.pos[k] + S.vel[k] * t;
            P.prm[k] = S.prm[k];
        }
#pragma unroll
        for (int k = 0; k < 4; k++) P.q[k] = S.quat[k];
        double p1[3], p2[3], p3[3];
        world_pos(rest, pos, lo, f.x, false, p1);
        world_pos(rest, pos, lo, f.y, false, p2);
        world_pos(rest, pos, lo, f.z, false, p3);
        double e12 = 0, e13 = 0, e23 = 0, cen[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            e12 += (p2[k] - p1[k]) * (p2[k] - p1[k]);
            e13 += (p3[k] - p1[k]) * (p3[k] - p1[k]);
            e23 += (p3[k] - p2[k]) * (p3[k] - p2[k]);
            cen[k] = (p1[k] + p2[k] + p3[k]) / 3;
        }
        const double me = fmax(e12, fmax(e13, e23));
        const double cd = sdf_eval(P, cen);
        if (!(cd * cd > me)) {
            // Frank-Wolfe, 32 steps of size 2/(i+3), started at the closest vertex
            const double d1 = sdf_eval(P, p1), d2 = sdf_eval(P, p2), d3 = sdf_eval(P, p3);
            double x[3];
            const double* st = (d1 <= d2 && d1 <= d3) ? p1 : ((d2 <= d1 && d2 <= d3) ? p2 : p3);
#pragma unroll
            for (int k = 0; k < 3; k++) x[k] = st[k];
            for (int it = 0; it < 32; it++) {
                const double al = 2.0 / (it + 3);
                double g[3];
                sdf_grad(P, x, g);
                const double s1 = p1[0] * g[0] + p1[1] * g[1] + p1[2] * g[2];
                const double s2 = p2[0] * g[0] + p2[1] * g[1] + p2[2] * g[2];
                const double s3 = p3[0] * g[0] + p3[1] * g[1] + p3[2] * g[2];
                const double* sv = (s1 < s2 && s1 < s3) ? p1 : ((s2 < s1 && s2 < s3) ? p2 : p3);
#pragma unroll
                for (int k = 0; k < 3; k++) x[k] = x[k] + al * (sv[k] - x[k]);
            }
            const double dist = sdf_eval(P, x);
            if (dist <= 1e-4) {
                double v0[3], v1[3], v2[3];
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    v0[k] = p2[k] - p1[k];
                    v1[k] = p3[k] - p1[k];
                    v2[k] = x[k] - p1[k];
                }
                const double d00 = v0[0] * v0[0] + v0[1] * v0[1] + v0[2] * v0[2];
                const double d01 = v0[0] * v1[0] + v0[1] * v1[1] + v0[2] * v1[2];
                const double d11 = v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2];
                const double d20 = v2[0] * v0[0] + v2[1] * v0[1] + v2[2] * v0[2];
                const double d21 = v2[0] * v1[0] + v2[1] * v1[1] + v2[2] * v1[2];
                const double den = d00 * d11 - d01 * d01;
                const double bv = (d11 * d20 - d01 * d21) / den;
                const double bw = (d00 * d21 - d01 * d20) / den;
                double g[3];
                sdf_grad(P, x, g);
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    cp[3 * (size_t)j + k] = x[k] - g[k] * dist;
                    cn[3 * (size_t)j + k] = g[k];
                }
                cb[3 * (size_t)j] = 1 - bv - bw;
                cb[3 * (size_t)j + 1] = bv;
                cb[3 * (size_t)j + 2] = bw;
                hit = 1;
            }
        }
    }
    flag[j] = hit;
}

// Ordered compaction of the hit flags in three small kernels:
//   count   per 256-face block, hits counted with wave ballots
//   scan    one workgroup turns the block counts into block offsets
//   scatter each hit's position = block offset + hits before it in the block
// Result: cpos[j] = number of hits before face slot j, cpos[nf] = total,
// clist[cpos[j]] = j for every hit (face order preserved).
constexpr int kScan = 1024;
constexpr int kWaves = kBlock / 64;

__global__ __launch_bounds__(kBlock) void k_count_hits(const int* __restrict__ flag, int nf, int* __restrict__ bsum)
{
    __shared__ int ws[kWaves];
    const int j = blockIdx.x * kBlock + threadIdx.x;
    const unsigned long long m = __ballot(j < nf && flag[j] != 0);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < kWaves; w++) t += ws[w];
        bsum[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(kScan) void k_scan_blocks(int* __restrict__ bsum, int nb, int* __restrict__ cpos, int nf)
{
    __shared__ int sh[kScan];
    const int chunk = (nb + kScan - 1) / kScan;
    const int b = threadIdx.x * chunk;
    const int e = min(nb, b + chunk);
    int cnt = 0;
    for (int j = b; j < e; j++) cnt += bsum[j];
    sh[threadIdx.x] = cnt;
    __syncthreads();
    for (int off = 1; off < kScan; off <<= 1) {  // Hillis-Steele inclusive scan
        const int v = (int)threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
        __syncthreads();
        sh[threadIdx.x] += v;
        __syncthreads();
    }
    int run = sh[threadIdx.x] - cnt;
    for (int j = b; j < e; j++) {
        const int v = bsum[j];
        bsum[j] = run;
        run += v;
    }
    if (threadIdx.x == kScan - 1) cpos[nf] = sh[kScan - 1];
}

__global__ __launch_bounds__(kBlock) void k_scatter_hits(const int* __restrict__ flag, int nf,
                                                        const int* __restrict__ boff, int* __restrict__ cpos,
                                                        int* __restrict__ clist)
{
    __shared__ int ws[kWaves];
    const int j = blockIdx.x * kBlock + threadIdx.x;
    const bool f = j < nf && flag[j] != 0;
    const unsigned long long m = __ballot(f);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) ws[wv] = __popcll(m);
    __syncthreads();
    int base = boff[blockIdx.x];
    for (int w = 0; w < wv; w++) base += ws[w];
    const int p = base + __popcll(m & ((1ull << lane) - 1ull));
    if (j < nf) {
        cpos[j] = p;
        if (f) clist[p] = j;
    }
}

// Static contact projection in FP64 (alpha = 0, inequality):
// ConstraintProjector.hpp:111-172, StaticDeformableCollisionConstraint.cpp:20-55
__device__ __forceinline__ bool contact_dx(const double* const x[3], const float* w, const double* cp,
                                           const double* cn, const double* cb, double* lam, double* dx)
{
    double a[3];
#pragma unroll
    for (int k = 0; k < 3; k++) a[k] = cb[0] * x[0][k] + cb[1] * x[1][k] + cb[2] * x[2][k];
    const double C = cn[0] * (a[0] - cp[0]) + cn[1] * (a[1] - cp[1]) + cn[2] * (a[2] - cp[2]);
    if (C > 0) return false;
    double L = 0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const double gx = cb[i] * cn[0], gy = cb[i] * cn[1], gz = cb[i] * cn[2];
        L += (double)w[i] * (gx * gx + gy * gy + gz * gz);
    }
    const double dl = -C / L;
    *lam += dl;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) dx[3 * i + k] = (double)w[i] * (cb[i] * cn[k]) * dl;
    return true;
}

// One contact of the Gauss-Seidel pass: project it from the current positions
// and apply its correction to the face's three vertices.
template <bool SPLIT>
__device__ __forceinline__ void contact_gs_one(bool first, int j, const int4* __restrict__ faces,
                                               const double* __restrict__ rest, float4* __restrict__ pos,
                                               float4* __restrict__ dsep, const float4* __restrict__ dbase,
                                               const float4* __restrict__ lo, const double* __restrict__ cp,
                                               const double* __restrict__ cn, const double* __restrict__ cb,
                                               double* __restrict__ clam)
{
    const int4 f = faces[j];
    double x0[3], x1[3], x2[3];
    if constexpr (SPLIT) {
        // world_pos with the delta read from the split array, in the same order
        const int vv[3] = { f.x, f.y, f.z };
        double* xx[3] = { x0, x1, x2 };
#pragma unroll
        for (int m = 0; m < 3; m++) {
            const float4 b = dbase[vv[m]];
            const float4 l = lo[vv[m]];
            double u[3] = { (double)b.x + (double)l.x, (double)b.y + (double)l.y, (double)b.z + (double)l.z };
            const float4 d = dsep[vv[m]];
            u[0] += (double)d.x;
            u[1] += (double)d.y;
            u[2] += (double)d.z;
#pragma unroll
            for (int k = 0; k < 3; k++) xx[m][k] = rest[3 * (size_t)vv[m] + k] + u[k];
        }
    } else {
        world_pos(rest, pos, lo, f.x, true, x0);
        world_pos(rest, pos, lo, f.y, true, x1);
        world_pos(rest, pos, lo, f.z, true, x2);
    }
    const double* xs[3] = { x0, x1, x2 };
    float w[3];
    if constexpr (SPLIT) {
        w[0] = dbase[f.x].w;
        w[1] = dbase[f.y].w;
        w[2] = dbase[f.z].w;
    } else {
        w[0] = pos[2 * (size_t)f.x].w;
        w[1] = pos[2 * (size_t)f.y].w;
        w[2] = pos[2 * (size_t)f.z].w;
    }
    double lam = first ? 0.0 : clam[j];
    double dx[9];
    if (contact_dx(xs, w, cp + 3 * (size_t)j, cn + 3 * (size_t)j, cb + 3 * (size_t)j, &lam, dx)) {
        const int vid[3] = { f.x, f.y, f.z };
#pragma unroll
        for (int i = 0; i < 3; i++) {
            float4 d = *delta_at<SPLIT>(pos, dsep, vid[i]);
            d.x = (float)((double)d.x + dx[3 * i]);
            d.y = (float)((double)d.y + dx[3 * i + 1]);
            d.z = (float)((double)d.z + dx[3 * i + 2]);
            *delta_at<SPLIT>(pos, dsep, vid[i]) = d;
        }
    }
    clam[j] = lam;
}

// Workgroup i projects the active contacts of instance i colour by colour
// (faces of one colour share no vertex; a barrier separates the colours), in
// face order within a colour; instances share no vertex, so their workgroups
// run side by side.  The contact range of every face colour is loaded up
// front, kBlock colours at a time.  The instance's contacts are numbered in
// colour order, and thread t owns contact t: everything that stays fixed
// during the solve (its face, rest positions, base + low displacement, inverse
// masses, contact point, normal, barycentrics, multiplier) is loaded once, so a
// colour costs one round trip for the in-step motion of its vertices.  Contacts
// past the first kBlock of an instance are loaded as they are projected.
template <bool FIRST, bool SPLIT>
__global__ __launch_bounds__(kBlock) void k_contacts_gs(const int4* __restrict__ faces,
                                                       const int* __restrict__ fci_off, int ncol, int ninst,
                                                       const int* __restrict__ cpos, const int* __restrict__ clist,
                                                       const double* __restrict__ rest, float4* __restrict__ pos,
                                                       float4* __restrict__ dsep, const float4* __restrict__ dbase,
                                                       const float4* __restrict__ lo,
                                                       const double* __restrict__ cp, const double* __restrict__ cn,
                                                       const double* __restrict__ cb, double* __restrict__ clam)
{
    __shared__ int rng[2 * kBlock];  // per colour of the chunk: the instance's range in clist
    __shared__ int loc[kBlock + 1];  // per colour: number of the instance's first contact of that colour
    const int tid = threadIdx.x, inst = blockIdx.x;
    for (int c0 = 0; c0 < ncol; c0 += kBlock) {
        const int nc = min(kBlock, ncol - c0);
        if (tid < nc) {
            const int* o = fci_off + (size_t)(c0 + tid) * (ninst + 1) + inst;
            rng[2 * tid] = cpos[o[0]];
            rng[2 * tid + 1] = cpos[o[1]];
        }
        __syncthreads();
        if (tid == 0) {
            int acc = 0;
            for (int c = 0; c < nc; c++) {
                loc[c] = acc;
                acc += rng[2 * c + 1] - rng[2 * c];
            }
            loc[nc] = acc;
        }
        __syncthreads();
        // this thread's contact: its colour (loc is non-decreasing) and fixed data
        int mc = -1, j = 0;
        int vid[3] = { 0, 0, 0 };
        double xr[3][3], sbl[3][3], ccp[3], ccn[3], ccb[3], lam = 0.0;
        float w[3] = { 0.f, 0.f, 0.f };
        if (tid < loc[nc]) {
            int lo_c = 0, hi_c = nc - 1;
            while (lo_c < hi_c) {
                const int mid = (lo_c + hi_c + 1) >> 1;
                if (loc[mid] <= tid) lo_c = mid;
                else hi_c = mid - 1;
            }
            mc = lo_c;
            j = clist[rng[2 * mc] + (tid - loc[mc])];
            const int4 f = faces[j];
            vid[0] = f.x;
            vid[1] = f.y;
            vid[2] = f.z;
#pragma unroll
            for (int m = 0; m < 3; m++) {
                const float4 bv = SPLIT ? dbase[vid[m]] : pos[2 * (size_t)vid[m]];
                const float4 l = lo[vid[m]];
                sbl[m][0] = (double)bv.x + (double)l.x;
                sbl[m][1] = (double)bv.y + (double)l.y;
                sbl[m][2] = (double)bv.z + (double)l.z;
                w[m] = bv.w;
#pragma unroll
                for (int k = 0; k < 3; k++) xr[m][k] = rest[3 * (size_t)vid[m] + k];
            }
#pragma unroll
            for (int k = 0; k < 3; k++) {
                ccp[k] = cp[3 * (size_t)j + k];
                ccn[k] = cn[3 * (size_t)j + k];
                ccb[k] = cb[3 * (size_t)j + k];
            }
            if (!FIRST) lam = clam[j];
        }
        for (int c = 0; c < nc; c++) {
            const int cs = rng[2 * c], ce = rng[2 * c + 1];
            if (cs == ce) continue;  // the same for every thread
            if (mc == c) {
                // world position = rest + ((base + low) + in-step motion), as world_pos
                float4 d[3];
                double x[3][3];
#pragma unroll
                for (int m = 0; m < 3; m++) {
                    d[m] = *delta_at<SPLIT>(pos, dsep, vid[m]);
                    const double u[3] = { sbl[m][0] + (double)d[m].x, sbl[m][1] + (double)d[m].y,
                                          sbl[m][2] + (double)d[m].z };
#pragma unroll
                    for (int k = 0; k < 3; k++) x[m][k] = xr[m][k] + u[k];
                }
                const double* xs[3] = { x[0], x[1], x[2] };
                double dx[9];
                if (contact_dx(xs, w, ccp, ccn, ccb, &lam, dx)) {
#pragma unroll
                    for (int m = 0; m < 3; m++) {
                        float4 dn = d[m];
                        dn.x = (float)((double)dn.x + dx[3 * m]);
                        dn.y = (float)((double)dn.y + dx[3 * m + 1]);
                        dn.z = (float)((double)dn.z + dx[3 * m + 2]);
                        *delta_at<SPLIT>(pos, dsep, vid[m]) = dn;
                    }
                }
            }
            for (int q = loc[c] + tid; q < loc[c + 1]; q += kBlock)
                if (q >= kBlock)
                    contact_gs_one<SPLIT>(FIRST, clist[cs + (q - loc[c])], faces, rest, pos, dsep, dbase, lo, cp, cn,
                                          cb, clam);
            __syncthreads();
        }
        if (mc >= 0) clam[j] = lam;
        __syncthreads();  // rng and loc are rewritten by the next chunk
    }
}

template <bool FIRST>
__global__ __launch_bounds__(kBlock) void k_contacts_jacobi(const int4* __restrict__ faces, int nf,
                                                           const int* __restrict__ cpos, const int* __restrict__ clist,
                                                           const double* __restrict__ rest, const float4* __restrict__ pos,
                                                           const float4* __restrict__ lo,
                                                           float* __restrict__ acc, const double* __restrict__ cp,
                                                           const double* __restrict__ cn, const double* __restrict__ cb,
                                                           double* __restrict__ clam)
{
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= cpos[nf]) return;
    const int j = clist[k];
    const int4 f = faces[j];
    double x0[3], x1[3], x2[3];
    world_pos(rest, pos, lo, f.x, true, x0);
    world_pos(rest, pos, lo, f.y, true, x1);
    world_pos(rest, pos, lo, f.z, true, x2);
    const double* xs[3] = { x0, x1, x2 };
    const float w[3] = { pos[2 * (size_t)f.x].w, pos[2 * (size_t)f.y].w, pos[2 * (size_t)f.z].w };
    double lam = FIRST ? 0.0 : clam[j];
    double dx[9];
    if (contact_dx(xs, w, cp + 3 * (size_t)j, cn + 3 * (size_t)j, cb + 3 * (size_t)j, &lam, dx)) {
        const int vid[3] = { f.x, f.y, f.z };
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int kk = 0; kk < 3; kk++) unsafeAtomicAdd(acc + 4 * (size_t)vid[i] + kk, (float)dx[3 * i + kk]);
    }
    clam[j] = lam;
}

// Fixed-order variant for the deterministic Jacobi mode: contact j leaves its
// nine corrections (zeros when inactive) in its face slot of cdx instead of
// adding them to the accumulator; k_contact_gather then adds them per vertex
// in ascending face-slot order.
template <bool FIRST>
__global__ __launch_bounds__(kBlock) void k_contacts_jacobi_det(const int4* __restrict__ faces, int nf,
                                                               const int* __restrict__ cpos,
                                                               const int* __restrict__ clist,
                                                               const double* __restrict__ rest,
                                                               const float4* __restrict__ pos,
                                                               const float4* __restrict__ lo, float* __restrict__ cdx,
                                                               const double* __restrict__ cp,
                                                               const double* __restrict__ cn,
                                                               const double* __restrict__ cb, double* __restrict__ clam)
{
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= cpos[nf]) return;
    const int j = clist[k];
    const int4 f = faces[j];
    double x0[3], x1[3], x2[3];
    world_pos(rest, pos, lo, f.x, true, x0);
    world_pos(rest, pos, lo, f.y, true, x1);
    world_pos(rest, pos, lo, f.z, true, x2);
    const double* xs[3] = { x0, x1, x2 };
    const float w[3] = { pos[2 * (size_t)f.x].w, pos[2 * (size_t)f.y].w, pos[2 * (size_t)f.z].w };
    double lam = FIRST ? 0.0 : clam[j];
    double dx[9];
    const bool act = contact_dx(xs, w, cp + 3 * (size_t)j, cn + 3 * (size_t)j, cb + 3 * (size_t)j, &lam, dx);
#pragma unroll
    for (int i = 0; i < 9; i++) cdx[9 * (size_t)j + i] = act ? (float)dx[i] : 0.f;
    clam[j] = lam;
}
#pragma clang fp contract(fast)

// Per vertex, the contact corrections of its faces (vf_off / vf: face slot * 3
// + corner, ascending face slot) that hold a contact this step, summed in that
// order and added to the accumulator.
__global__ __launch_bounds__(kBlock) void k_contact_gather(int nv, const int* __restrict__ vf_off,
                                                          const int* __restrict__ vf, const int* __restrict__ flag,
                                                          const float* __restrict__ cdx, float* __restrict__ acc)
{
    const int v = blockIdx.x * kBlock + threadIdx.x;
    if (v >= nv) return;
    const int b = vf_off[v], e = vf_off[v + 1];
    if (b == e) return;
    float s0 = 0.f, s1 = 0.f, s2 = 0.f;
    bool any = false;
    for (int k = b; k < e; k++) {
        const int jc = vf[k], j = jc / 3, c = jc - 3 * j;
        if (!flag[j]) continue;
        const float* d = cdx + 9 * (size_t)j + 3 * c;
        s0 += d[0];
        s1 += d[1];
        s2 += d[2];
        any = true;
    }
    if (any) {
        acc[4 * (size_t)v] += s0;
        acc[4 * (size_t)v + 1] += s1;
        acc[4 * (size_t)v + 2] += s2;
    }
}

// ---------------------------------------------------------------------------
// Fused patch Jacobi (fixed-order accumulation), one launch per iteration.
// A shared vertex's value after the previous iteration is formed while it is
// staged: its value before that iteration (brd, kept by the patch that holds
// the vertex's first slot) plus the previous iteration's partial sums in slot
// order, the same float operations as the separate shared-vertex pass.  The
// partial sums and those values alternate between two buffers by iteration
// parity, so no patch reads what another writes in the same launch.  Contacts
// are projected by the patch that holds their face, from its staged positions,
// face by face, and summed after that patch's element corrections.
// k_jacobi_fused_final writes the shared vertices once the last iteration is done.
// ---------------------------------------------------------------------------
constexpr int kFusedSlots = 4;  // partial slots of a shared vertex loaded together while staging

template <bool FIRST, bool CON>
__global__ __launch_bounds__(kPatch) void k_jacobi_fused(TetArrays ta, FusedPatchArgs A)
{
    extern __shared__ float4 smem[];
    const int maxl = A.maxl;
    float4* lb = smem;                                   // base, inverse mass
    float* ld = reinterpret_cast<float*>(smem + maxl);   // in-step motion, 3 per local vertex
    float* red = ld + 3 * maxl;                          // 12 planes of kPatch element corrections
    uint32_t* linc = reinterpret_cast<uint32_t*>(red + 12 * kPatch);  // the patch's incidence list
    float* lcon = reinterpret_cast<float*>(linc + 2 * kPatch);        // CON: contact corrections, 3 per local vertex
    float* fdx = lcon + 3 * maxl;                                      // CON: one batch of faces' corrections, 9 each
    int* fact = reinterpret_cast<int*>(fdx + 9 * kPatch);              // CON: which faces of the batch are active
    const int tid = threadIdx.x;
    const int p = xcd_item(blockIdx.x, gridDim.x, A.xcd != 0);
    const int v0 = A.pvoff[p], nl = A.pvoff[p + 1] - v0;
    const int e = p * kPatch + tid;
    const bool has = e < A.nt;
    uint2 lc = make_uint2(0u, 0u);
    float q[9];
    float2 at = make_float2(0.f, 0.f);
    float gm = 0.f;
    float l[2] = { 0.0f, 0.0f };
    if (has) {
        lc = A.tloc[e];
        load_tet(ta, e, q, at, gm);
        if (!FIRST) {
            const float2 lv = ta.lam[e];
            l[0] = lv.x;
            l[1] = lv.y;
        }
    }
    {
        const uint2 en = reinterpret_cast<const uint2*>(A.inc + (size_t)p * 4 * kPatch)[tid];
        linc[2 * tid] = en.x;
        linc[2 * tid + 1] = en.y;
    }
    for (int i = tid; i < nl; i += kPatch) {
        const int g = A.pverts[v0 + i];
        lb[i] = A.pos[2 * (size_t)g];
        const int2 sh = A.pshare[v0 + i];
        float4 d;
        if (sh.x < 0 || A.first_iter) {
            d = A.pos[2 * (size_t)g + 1];
        } else {
            // the first kFusedSlots partials are loaded together (predicated), then
            // summed in slot order; rarer vertices with more slots loop over the rest
            const int s0 = sh.y & ((1 << 28) - 1), cnt = (int)((unsigned)sh.y >> 28);
            float4 v[kFusedSlots];
#pragma unroll
            for (int k = 0; k < kFusedSlots; k++) v[k] = k < cnt ? A.prd[s0 + k] : make_float4(0.f, 0.f, 0.f, 0.f);
            d = A.brd[sh.x];
            float a0 = 0.f, a1 = 0.f, a2 = 0.f;
#pragma unroll
            for (int k = 0; k < kFusedSlots; k++)
                if (k < cnt) {
                    a0 += v[k].x;
                    a1 += v[k].y;
                    a2 += v[k].z;
                }
            for (int k = kFusedSlots; k < cnt; k++) {
                const float4 u = A.prd[s0 + k];
                a0 += u.x;
                a1 += u.y;
                a2 += u.z;
            }
            d.x += a0;
            d.y += a1;
            d.z += a2;
        }
        // the patch holding the vertex's first slot keeps this value for the next iteration
        if (sh.x >= 0 && A.pslot[v0 + i] == (sh.y & ((1 << 28) - 1))) A.bwr[sh.x] = d;
        ld[3 * i] = d.x;
        ld[3 * i + 1] = d.y;
        ld[3 * i + 2] = d.z;
        if (CON) lcon[3 * i] = lcon[3 * i + 1] = lcon[3 * i + 2] = 0.f;
    }
    __syncthreads();
    if (has) {
        const int sl[4] = { (int)(lc.x & 0xffffu), (int)(lc.x >> 16), (int)(lc.y & 0xffffu), (int)(lc.y >> 16) };
        float bb[12], dd[12], w[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const float4 b = lb[sl[j]];
            bb[3 * j] = b.x;
            bb[3 * j + 1] = b.y;
            bb[3 * j + 2] = b.z;
            w[j] = b.w;
#pragma unroll
            for (int k = 0; k < 3; k++) dd[3 * j + k] = ld[3 * sl[j] + k];
        }
        float dx[12];
        project_tet_split<float>(bb, dd, w, q, at.x, at.y, gm, l, dx);
        ta.lam[e] = make_float2(l[0], l[1]);
#pragma unroll
        for (int k = 0; k < 12; k++) red[k * kPatch + tid] = dx[k];
    }
    if (CON) {
        // the patch's surface faces, a batch of kPatch at a time: each thread
        // projects one face in FP64 on the staged state, then thread 0 adds the
        // batch's corrections in face order (a fixed summation order)
        const int f0 = A.pf_off[p], f1 = A.pf_off[p + 1];
        for (int fb = f0; fb < f1; fb += kPatch) {
            const int f = fb + tid;
            bool act = false;
            if (f < f1) {
                const int4 F = A.pf[f];  // (face, local slots of its three vertices)
                if (A.flag[F.x]) {
                    const int ls[3] = { F.y, F.z, F.w };
                    double x[3][3];
                    float w[3];
#pragma unroll
                    for (int m = 0; m < 3; m++) {
                        const int g = A.pverts[v0 + ls[m]];
                        const float4 b = lb[ls[m]];
                        const float4 lo = A.lo[g];
                        x[m][0] = A.rest[3 * (size_t)g] + (((double)b.x + (double)lo.x) + (double)ld[3 * ls[m]]);
                        x[m][1] = A.rest[3 * (size_t)g + 1] + (((double)b.y + (double)lo.y) + (double)ld[3 * ls[m] + 1]);
                        x[m][2] = A.rest[3 * (size_t)g + 2] + (((double)b.z + (double)lo.z) + (double)ld[3 * ls[m] + 2]);
                        w[m] = b.w;
                    }
                    const double* xs[3] = { x[0], x[1], x[2] };
                    double lam = FIRST ? 0.0 : A.clam[F.x];
                    double cdx[9];
                    act = contact_dx(xs, w, A.cp + 3 * (size_t)F.x, A.cn + 3 * (size_t)F.x, A.cb + 3 * (size_t)F.x,
                                     &lam, cdx);
                    if (act)
                        for (int k = 0; k < 9; k++) fdx[9 * tid + k] = (float)cdx[k];
                    A.clam[F.x] = lam;
                }
            }
            fact[tid] = act ? 1 : 0;
            __syncthreads();
            if (tid == 0)
                for (int t = 0; t < min(kPatch, f1 - fb); t++) {
                    if (!fact[t]) continue;
                    const int4 F = A.pf[fb + t];
                    const int ls[3] = { F.y, F.z, F.w };
                    for (int m = 0; m < 3; m++)
                        for (int k = 0; k < 3; k++) lcon[3 * ls[m] + k] += fdx[9 * t + 3 * m + k];
                }
            __syncthreads();
        }
    }
    __syncthreads();
    const int ioff = v0 + p;  // this patch's nl+1 incidence offsets
    for (int i = tid; i < nl; i += kPatch) {
        float a0 = 0.f, a1 = 0.f, a2 = 0.f;
        const int s0 = A.inc_off[ioff + i], s1 = A.inc_off[ioff + i + 1];
        gather_incidence(reinterpret_cast<const uint16_t*>(linc), red, s0, s1, a0, a1, a2);
        if (CON) {
            a0 += lcon[3 * i];
            a1 += lcon[3 * i + 1];
            a2 += lcon[3 * i + 2];
        }
        const int g = A.pverts[v0 + i];
        const int slot = A.pslot[v0 + i];
        if (slot < 0)
            A.pos[2 * (size_t)g + 1] = make_float4(ld[3 * i] + a0, ld[3 * i + 1] + a1, ld[3 * i + 2] + a2, 0.f);
        else
            A.pwr[slot] = make_float4(a0, a1, a2, 0.f);
    }
}

// After the last fused iteration: shared vertex j = its value before that
// iteration + that iteration's partial sums in slot order.
__global__ __launch_bounds__(kBlock) void k_jacobi_fused_final(const int* __restrict__ shared,
                                                              const int* __restrict__ soff,
                                                              const float4* __restrict__ prd,
                                                              const float4* __restrict__ brd,
                                                              float4* __restrict__ pos, int ns)
{
    const int j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= ns) return;
    float a0 = 0.f, a1 = 0.f, a2 = 0.f;
    for (int k = soff[j]; k < soff[j + 1]; k++) {
        const float4 v = prd[k];
        a0 += v.x;
        a1 += v.y;
        a2 += v.z;
    }
    float4 d = brd[j];
    d.x += a0;
    d.y += a1;
    d.z += a2;
    pos[2 * (size_t)shared[j] + 1] = d;
}

// Attachments (AttachmentConstraint.hpp:31-71 with ConstraintProjector.hpp:111-172,
// alpha = 0): vertex v pulled to the point p, C = |x - p|, gradient
// (x - p) / |x - p|; no correction below |x - p| = 1e-12, where the constraint
// is met (the reference's 1e-12 substitute gradient, AttachmentConstraint.hpp:76-81,
// would move the vertex by ~3e11 |x - p|).  FP64 on the FP32 state, like the
// contacts.  Attached vertices are distinct, so one thread per attachment is
// the sequential order.
// ---------------------------------------------------------------------------
#pragma clang fp contract(off)
__device__ __forceinline__ void attach_dx(const double* x, const double* p, float w, double* lam, double* dx)
{
    const double d[3] = { x[0] - p[0], x[1] - p[1], x[2] - p[2] };
    const double dist = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (dist < 1e-12) {
        dx[0] = dx[1] = dx[2] = 0.0;
        return;
    }
    double g[3];
    for (int k = 0; k < 3; k++) g[k] = d[k] / dist;
    const double lhs = (double)w * (g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
    const double dlam = -dist / lhs;
    *lam += dlam;
    for (int k = 0; k < 3; k++) dx[k] = (double)w * g[k] * dlam;
}
#pragma clang fp contract(fast)

// Attachments a in [a0, n) by stride, each applied in place.
__device__ __forceinline__ void attach_gs_pass(bool first, int a0, int stride, const int* __restrict__ av,
                                               const double* __restrict__ ap, double* __restrict__ alam,
                                               const double* __restrict__ rest, float4* __restrict__ pos,
                                               const float4* __restrict__ lo, int n)
{
    for (int a = a0; a < n; a += stride) {
        const int v = av[a];
        double x[3], dx[3];
        world_pos(rest, pos, lo, v, true, x);
        double lam = first ? 0.0 : alam[a];
        attach_dx(x, ap + 3 * (size_t)a, pos[2 * (size_t)v].w, &lam, dx);
        alam[a] = lam;
        float4 d = pos[2 * (size_t)v + 1];
        d.x += (float)dx[0];
        d.y += (float)dx[1];
        d.z += (float)dx[2];
        pos[2 * (size_t)v + 1] = d;
    }
}

template <bool FIRST>
__global__ __launch_bounds__(kBlock) void k_attach_gs(const int* __restrict__ av, const double* __restrict__ ap,
                                                     double* __restrict__ alam, const double* __restrict__ rest,
                                                     float4* __restrict__ pos, const float4* __restrict__ lo, int n)
{
    const int a = blockIdx.x * kBlock + threadIdx.x;
    if (a >= n) return;
    attach_gs_pass(FIRST, a, n, av, ap, alam, rest, pos, lo, n);
}

template <bool FIRST>
__global__ __launch_bounds__(kBlock) void k_attach_jacobi(const int* __restrict__ av, const double* __restrict__ ap,
                                                         double* __restrict__ alam, const double* __restrict__ rest,
                                                         const float4* __restrict__ pos, const float4* __restrict__ lo,
                                                         float* __restrict__ acc, int n)
{
    const int a = blockIdx.x * kBlock + threadIdx.x;
    if (a >= n) return;
    const int v = av[a];
    double x[3], dx[3];
    world_pos(rest, pos, lo, v, true, x);
    double lam = FIRST ? 0.0 : alam[a];
    attach_dx(x, ap + 3 * (size_t)a, pos[2 * (size_t)v].w, &lam, dx);
    alam[a] = lam;
    // attached vertices are distinct (xpbd_set_attachments checks it): one writer per slot
    for (int k = 0; k < 3; k++) acc[4 * (size_t)v + k] += (float)dx[k];
}

// ---------------------------------------------------------------------------
// Local collision iterations (XPBDMeshObject.cpp:476-483, 764-825).  After the
// global solve, mult[p] counts the (contact with non-zero multiplier, face
// vertex) pairs whose vertex belongs to the element at solver position p; a
// local pass projects every such element mult[p] times in a row (the
// reference's list holds it that many times), in the global solve's order,
// without resetting multipliers, then all contacts.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_local_count(const int4* __restrict__ faces, int nf,
                                                       const int* __restrict__ cpos, const int* __restrict__ clist,
                                                       const double* __restrict__ clam, const int* __restrict__ vt_off,
                                                       const int* __restrict__ vt_pos, int* __restrict__ mult)
{
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= cpos[nf]) return;
    const int j = clist[k];
    if (clam[j] == 0.0) return;
    const int4 f = faces[j];
    const int vid[3] = { f.x, f.y, f.z };
    for (int i = 0; i < 3; i++)
        for (int p = vt_off[vid[i]]; p < vt_off[vid[i] + 1]; p++) atomicAdd(mult + vt_pos[p], 1);
}

// Gauss-Seidel: project element e m times, each projection seeing the last one's
// positions and multipliers, then write its four vertices back.
__device__ __forceinline__ void project_repeated(const TetArrays& ta, int e, int m, float4* pos)
{
    const int4 t = ta.tet[e];
    Gather g;
    gather4(pos, t, g);
    float q[9];
    float2 at;
    float gm;
    load_tet(ta, e, q, at, gm);
    const float2 lv = ta.lam[e];
    float l[2] = { lv.x, lv.y };
    for (int r = 0; r < m; r++) {
        float dx[12];
        project_tet_split<float>(g.b, g.d, g.w, q, at.x, at.y, gm, l, dx);
#pragma unroll
        for (int k = 0; k < 12; k++) g.d[k] += dx[k];
    }
    ta.lam[e] = make_float2(l[0], l[1]);
    const int vid[4] = { t.x, t.y, t.z, t.w };
#pragma unroll
    for (int i = 0; i < 4; i++) pos[2 * (size_t)vid[i] + 1] = make_float4(g.d[3 * i], g.d[3 * i + 1], g.d[3 * i + 2], 0.f);
}

// one thread per element of a colour (elements of a colour share no vertex)
__global__ __launch_bounds__(kBlock) void k_gs_local_color(TetArrays ta, const int* __restrict__ mult,
                                                          float4* __restrict__ pos, int begin, int end)
{
    const int e = begin + blockIdx.x * kBlock + threadIdx.x;
    if (e >= end) return;
    const int m = mult[e];
    if (m > 0) project_repeated(ta, e, m, pos);
}

// one thread per cluster of a colour, its elements in order (clusters of a
// colour share no vertex; a thread re-reads what it wrote itself)
__global__ __launch_bounds__(kBlock) void k_gs_local_cluster(TetArrays ta, const int* __restrict__ cstart,
                                                            const int* __restrict__ mult, float4* __restrict__ pos,
                                                            int cbeg, int cend)
{
    const int c = cbeg + blockIdx.x * kBlock + threadIdx.x;
    if (c >= cend) return;
    const int e1 = cstart[c + 1];
    for (int e = cstart[c]; e < e1; e++) {
        const int m = mult[e];
        if (m > 0) project_repeated(ta, e, m, pos);
    }
}

// Jacobi: the m projections of element e all read the pass's starting positions
// (the multiplier advances between them); their sum goes into acc.
// Fixed-order variant: each counted element leaves its summed corrections in
// its solver position's slot of edx; k_local_gather adds them per vertex.
__global__ __launch_bounds__(kBlock) void k_jacobi_local_det(TetArrays ta, const int* __restrict__ mult,
                                                            const float4* __restrict__ pos, float* __restrict__ edx,
                                                            int nt)
{
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= nt) return;
    const int m = mult[e];
    if (m == 0) return;
    const int4 t = ta.tet[e];
    Gather g;
    gather4(pos, t, g);
    float q[9];
    float2 at;
    float gm;
    load_tet(ta, e, q, at, gm);
    const float2 lv = ta.lam[e];
    float l[2] = { lv.x, lv.y };
    float sum[12] = {};
    for (int r = 0; r < m; r++) {
        float dx[12];
        project_tet_split<float>(g.b, g.d, g.w, q, at.x, at.y, gm, l, dx);
#pragma unroll
        for (int k = 0; k < 12; k++) sum[k] += dx[k];
    }
    ta.lam[e] = make_float2(l[0], l[1]);
#pragma unroll
    for (int k = 0; k < 12; k++) edx[12 * (size_t)e + k] = sum[k];
}

// Per vertex, the local-pass corrections of its counted elements in ascending
// solver position (vt_off / vt_pos), added to the accumulator.
__global__ __launch_bounds__(kBlock) void k_local_gather(TetArrays ta, int nv, const int* __restrict__ vt_off,
                                                        const int* __restrict__ vt_pos, const int* __restrict__ mult,
                                                        const float* __restrict__ edx, float* __restrict__ acc)
{
    const int v = blockIdx.x * kBlock + threadIdx.x;
    if (v >= nv) return;
    float s0 = 0.f, s1 = 0.f, s2 = 0.f;
    bool any = false;
    for (int k = vt_off[v]; k < vt_off[v + 1]; k++) {
        const int p = vt_pos[k];
        if (mult[p] == 0) continue;
        const int4 t = ta.tet[p];
        const int c = t.x == v ? 0 : t.y == v ? 1 : t.z == v ? 2 : 3;
        const float* d = edx + 12 * (size_t)p + 3 * c;
        s0 += d[0];
        s1 += d[1];
        s2 += d[2];
        any = true;
    }
    if (any) {
        acc[4 * (size_t)v] += s0;
        acc[4 * (size_t)v + 1] += s1;
        acc[4 * (size_t)v + 2] += s2;
    }
}

__global__ __launch_bounds__(kBlock) void k_jacobi_local(TetArrays ta, const int* __restrict__ mult,
                                                        const float4* __restrict__ pos, float* __restrict__ acc, int nt)
{
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= nt) return;
    const int m = mult[e];
    if (m == 0) return;
    const int4 t = ta.tet[e];
    Gather g;
    gather4(pos, t, g);
    float q[9];
    float2 at;
    float gm;
    load_tet(ta, e, q, at, gm);
    const float2 lv = ta.lam[e];
    float l[2] = { lv.x, lv.y };
    float sum[12] = {};
    for (int r = 0; r < m; r++) {
        float dx[12];
        project_tet_split<float>(g.b, g.d, g.w, q, at.x, at.y, gm, l, dx);
#pragma unroll
        for (int k = 0; k < 12; k++) sum[k] += dx[k];
    }
    ta.lam[e] = make_float2(l[0], l[1]);
    const int vid[4] = { t.x, t.y, t.z, t.w };
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) unsafeAtomicAdd(acc + 4 * (size_t)vid[i] + k, sum[3 * i + k]);
}

inline unsigned nblk(long long n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

hipError_t launch_attach_gs(hipStream_t s, const int* av, const double* ap, double* alam, const double* rest,
                            float4* pos, const float4* lo, int n, bool first)
{
    if (n <= 0) return hipSuccess;
    if (first) hipLaunchKernelGGL(k_attach_gs<true>, dim3(nblk(n, kBlock)), dim3(kBlock), 0, s, av, ap, alam, rest, pos, lo, n);
    else hipLaunchKernelGGL(k_attach_gs<false>, dim3(nblk(n, kBlock)), dim3(kBlock), 0, s, av, ap, alam, rest, pos, lo, n);
    return hipGetLastError();
}

hipError_t launch_attach_jacobi(hipStream_t s, const int* av, const double* ap, double* alam, const double* rest,
                                const float4* pos, const float4* lo, float* acc, int n, bool first)
{
    if (n <= 0) return hipSuccess;
    if (first)
        hipLaunchKernelGGL(k_attach_jacobi<true>, dim3(nblk(n, kBlock)), dim3(kBlock), 0, s, av, ap, alam, rest, pos, lo,
                           acc, n);
    else
        hipLaunchKernelGGL(k_attach_jacobi<false>, dim3(nblk(n, kBlock)), dim3(kBlock), 0, s, av, ap, alam, rest, pos,
                           lo, acc, n);
    return hipGetLastError();
}

hipError_t launch_local_count(hipStream_t s, const int4* faces, int nf, const int* cpos, const int* clist,
                              const double* clam, const int* vt_off, const int* vt_pos, int* mult)
{
    if (nf <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_local_count, dim3(nblk(nf, kBlock)), dim3(kBlock), 0, s, faces, nf, cpos, clist, clam, vt_off,
                       vt_pos, mult);
    return hipGetLastError();
}

hipError_t launch_gs_local_color(hipStream_t s, const TetArrays& ta, const int* mult, float4* pos, int begin, int end)
{
    if (end <= begin) return hipSuccess;
    hipLaunchKernelGGL(k_gs_local_color, dim3(nblk(end - begin, kBlock)), dim3(kBlock), 0, s, ta, mult, pos, begin, end);
    return hipGetLastError();
}

hipError_t launch_gs_local_cluster(hipStream_t s, const TetArrays& ta, const int* cstart, const int* mult, float4* pos,
                                   int cbeg, int cend)
{
    if (cend <= cbeg) return hipSuccess;
    hipLaunchKernelGGL(k_gs_local_cluster, dim3(nblk(cend - cbeg, kBlock)), dim3(kBlock), 0, s, ta, cstart, mult, pos,
                       cbeg, cend);
    return hipGetLastError();
}

hipError_t launch_jacobi_local(hipStream_t s, const TetArrays& ta, const int* mult, const float4* pos, float* acc, int nt)
{
    if (nt <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_jacobi_local, dim3(nblk(nt, kBlock)), dim3(kBlock), 0, s, ta, mult, pos, acc, nt);
    return hipGetLastError();
}

hipError_t launch_predict(hipStream_t s, float4* pos, const float4* vel, int nv, float dt, float dz)
{
    if (nv <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_predict, dim3(nblk(nv, kBlock)), dim3(kBlock), 0, s, pos, vel, nv, dt, dz);
    return hipGetLastError();
}

hipError_t launch_predict_first_order(hipStream_t s, float4* pos, const float* dz1, int nv)
{
    hipLaunchKernelGGL(k_predict_first_order, dim3(nblk(nv, kBlock)), dim3(kBlock), 0, s, pos, dz1, nv);
    return hipGetLastError();
}

hipError_t launch_finish(hipStream_t s, float4* pos, float4* lo, float4* vel, int nv, float dt,
                         unsigned long long* step_counter)
{
    hipLaunchKernelGGL(k_finish, dim3(nblk(nv > 0 ? nv : 1, kBlock)), dim3(kBlock), 0, s, pos, lo, vel, nv, dt,
                       step_counter);
    return hipGetLastError();
}

hipError_t launch_contacts_jacobi_det(hipStream_t s, const int4* faces, int nf, const int* cpos, const int* clist,
                                      const double* rest, const float4* pos, const float4* lo, float* cdx,
                                      const double* cp, const double* cn, const double* cb, double* clam, bool first)
{
    if (nf <= 0) return hipSuccess;
    const dim3 g(nblk(nf, kBlock));
    if (first)
        hipLaunchKernelGGL(k_contacts_jacobi_det<true>, g, dim3(kBlock), 0, s, faces, nf, cpos, clist, rest, pos, lo,
                           cdx, cp, cn, cb, clam);
    else
        hipLaunchKernelGGL(k_contacts_jacobi_det<false>, g, dim3(kBlock), 0, s, faces, nf, cpos, clist, rest, pos, lo,
                           cdx, cp, cn, cb, clam);
    return hipGetLastError();
}

hipError_t launch_contact_gather(hipStream_t s, int nv, const int* vf_off, const int* vf, const int* flag,
                                 const float* cdx, float* acc)
{
    if (nv <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_contact_gather, dim3(nblk(nv, kBlock)), dim3(kBlock), 0, s, nv, vf_off, vf, flag, cdx, acc);
    return hipGetLastError();
}

hipError_t launch_jacobi_local_det(hipStream_t s, const TetArrays& ta, const int* mult, const float4* pos, float* edx,
                                   int nt)
{
    if (nt <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_jacobi_local_det, dim3(nblk(nt, kBlock)), dim3(kBlock), 0, s, ta, mult, pos, edx, nt);
    return hipGetLastError();
}

hipError_t launch_local_gather(hipStream_t s, const TetArrays& ta, int nv, const int* vt_off, const int* vt_pos,
                               const int* mult, const float* edx, float* acc)
{
    if (nv <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_local_gather, dim3(nblk(nv, kBlock)), dim3(kBlock), 0, s, ta, nv, vt_off, vt_pos, mult, edx,
                       acc);
    return hipGetLastError();
}

hipError_t launch_commit(hipStream_t s, float4* pos, float4* lo, int nv)
{
    if (nv <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_commit, dim3(nblk(nv, kBlock)), dim3(kBlock), 0, s, pos, lo, nv);
    return hipGetLastError();
}

hipError_t launch_reset_listed(hipStream_t s, float2* lam, const int* mult, int nt)
{
    if (nt <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_reset_listed, dim3(nblk(nt, kBlock)), dim3(kBlock), 0, s, lam, mult, nt);
    return hipGetLastError();
}

hipError_t launch_lambda_start(hipStream_t s, const TetArrays& ta, const float4* pos, const float4* lo, int nt)
{
    if (nt <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_lambda_start, dim3(nblk(nt, kBlock)), dim3(kBlock), 0, s, ta, pos, lo, nt);
    return hipGetLastError();
}

hipError_t launch_gs_color(hipStream_t s, const TetArrays& ta, float4* pos, int begin, int end, bool first)
{
    if (end <= begin) return hipSuccess;
    const dim3 g(nblk(end - begin, kBlock));
    if (first) hipLaunchKernelGGL(k_gs_color<true>, g, dim3(kBlock), 0, s, ta, pos, begin, end);
    else hipLaunchKernelGGL(k_gs_color<false>, g, dim3(kBlock), 0, s, ta, pos, begin, end);
    return hipGetLastError();
}

hipError_t launch_gs_cluster(hipStream_t s, const TetArrays& ta, const int4* cverts, const int* cstart,
                             const uint32_t* tloc, float4* pos, int cbeg, int cend, bool first, bool xcd)
{
    if (cend <= cbeg) return hipSuccess;
    const dim3 g(nblk(cend - cbeg, kClusterBlock));
    if (first)
        hipLaunchKernelGGL(k_gs_cluster<true>, g, dim3(kClusterBlock), 0, s, ta, cverts, cstart, tloc, pos, cbeg, cend,
                           (int)xcd);
    else
        hipLaunchKernelGGL(k_gs_cluster<false>, g, dim3(kClusterBlock), 0, s, ta, cverts, cstart, tloc, pos, cbeg, cend,
                           (int)xcd);
    return hipGetLastError();
}

template <int BLOCK, int MT, bool LEAN, bool SPLIT>
static void launch_quad_shape(hipStream_t s, const TetArrays& ta, const int4* cverts, const int* cstart,
                              const uint32_t* tloc, float4* pos, float4* dsep, const float4* dbase, int cbeg,
                              int cend, bool first, bool xcd)
{
    const dim3 g(nblk(cend - cbeg, BLOCK / 4));
    if (first)
        hipLaunchKernelGGL((k_gs_cluster_quad<true, BLOCK, MT, LEAN, SPLIT>), g, dim3(BLOCK), 0, s, ta, cverts, cstart,
                           tloc, pos, dsep, dbase, cbeg, cend, (int)xcd);
    else
        hipLaunchKernelGGL((k_gs_cluster_quad<false, BLOCK, MT, LEAN, SPLIT>), g, dim3(BLOCK), 0, s, ta, cverts,
                           cstart, tloc, pos, dsep, dbase, cbeg, cend, (int)xcd);
}

template <int MT, bool LEAN, bool SPLIT>
static bool launch_quad_mt(hipStream_t s, const TetArrays& ta, const int4* cverts, const int* cstart,
                           const uint32_t* tloc, float4* pos, float4* dsep, const float4* dbase, int cbeg, int cend,
                           bool first, int block, bool xcd)
{
    if (block == 64)
        launch_quad_shape<64, MT, LEAN, SPLIT>(s, ta, cverts, cstart, tloc, pos, dsep, dbase, cbeg, cend, first, xcd);
    else if (block == 128)
        launch_quad_shape<128, MT, LEAN, SPLIT>(s, ta, cverts, cstart, tloc, pos, dsep, dbase, cbeg, cend, first, xcd);
    else if (block == 256)
        launch_quad_shape<256, MT, LEAN, SPLIT>(s, ta, cverts, cstart, tloc, pos, dsep, dbase, cbeg, cend, first, xcd);
    else return false;
    return true;
}

template <int MT, bool SPLIT>
static bool launch_quad_lean(hipStream_t s, const TetArrays& ta, const int4* cverts, const int* cstart,
                             const uint32_t* tloc, float4* pos, float4* dsep, const float4* dbase, int cbeg, int cend,
                             bool first, int block, bool xcd)
{
    // a negative shape selects the LDS-lean form
    return block < 0 ? launch_quad_mt<MT, true, SPLIT>(s, ta, cverts, cstart, tloc, pos, dsep, dbase, cbeg, cend,
                                                       first, -block, xcd)
                     : launch_quad_mt<MT, false, SPLIT>(s, ta, cverts, cstart, tloc, pos, dsep, dbase, cbeg, cend,
                                                        first, block, xcd);
}

int quad_shape(int block)
{
    const int b = block < 0 ? -block : block;
    const int shape = b > kQuadSplit ? b - kQuadSplit : b;
    return block < 0 ? -shape : shape;
}

bool quad_split(int block)
{
    return (block < 0 ? -block : block) > kQuadSplit;
}

bool quad_block_supported(int block)
{
    const int s = quad_shape(block);
    const int b = s < 0 ? -s : s;
    return b == 64 || b == 128 || b == 256;
}

hipError_t launch_gs_cluster_quad(hipStream_t s, const TetArrays& ta, const int4* cverts, const int* cstart,
                                  const uint32_t* tloc, float4* pos, int cbeg, int cend, bool first, int block,
                                  bool xcd, int max_tets, float4* dsep, const float4* dbase)
{
    if (cend <= cbeg) return hipSuccess;
    if (max_tets < 1 || max_tets > kClusterTets) return hipErrorInvalidValue;
    const int b = quad_shape(block);
    bool ok;
    if (dsep && dbase)
        ok = max_tets <= 6
                 ? launch_quad_lean<6, true>(s, ta, cverts, cstart, tloc, pos, dsep, dbase, cbeg, cend, first, b, xcd)
                 : launch_quad_lean<8, true>(s, ta, cverts, cstart, tloc, pos, dsep, dbase, cbeg, cend, first, b, xcd);
    else
        ok = max_tets <= 6 ? launch_quad_lean<6, false>(s, ta, cverts, cstart, tloc, pos, nullptr, nullptr, cbeg, cend,
                                                        first, b, xcd)
                           : launch_quad_lean<8, false>(s, ta, cverts, cstart, tloc, pos, nullptr, nullptr, cbeg, cend,
                                                        first, b, xcd);
    if (!ok) return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_split_deltas(hipStream_t s, const float4* pos, float4* dsep, int nv)
{
    if (nv <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_split_deltas, dim3(nblk(nv, kBlock)), dim3(kBlock), 0, s, pos, dsep, nv);
    return hipGetLastError();
}

hipError_t launch_join_deltas(hipStream_t s, float4* pos, const float4* dsep, int nv)
{
    if (nv <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_join_deltas, dim3(nblk(nv, kBlock)), dim3(kBlock), 0, s, pos, dsep, nv);
    return hipGetLastError();
}

hipError_t launch_jacobi_project(hipStream_t s, const TetArrays& ta, const float4* pos, float* acc, int nt,
                                 bool first)
{
    if (nt <= 0) return hipSuccess;
    const dim3 g(nblk(nt, kBlock));
    if (first) hipLaunchKernelGGL(k_jacobi_project<true>, g, dim3(kBlock), 0, s, ta, pos, acc, nt);
    else hipLaunchKernelGGL(k_jacobi_project<false>, g, dim3(kBlock), 0, s, ta, pos, acc, nt);
    return hipGetLastError();
}

size_t jacobi_patch_smem(int maxl, bool det)
{
    return (size_t)maxl * 16 + (size_t)maxl * 12 + (det ? (size_t)12 * kPatch * 4 + (size_t)4 * kPatch * 2 : (size_t)maxl * 12);
}

int jacobi_patch_size() { return kPatch; }

hipError_t launch_jacobi_patch(hipStream_t s, const TetArrays& ta, const PatchArrays& pa, float4* pos, float4* cacc,
                               int nt, bool first, bool det, bool con, bool xcd)
{
    if (pa.npatch <= 0) return hipSuccess;
    const size_t sm = jacobi_patch_smem(pa.maxl, det);
#define XPBD_PATCH_LAUNCH(F, D, Cn)                                                                        \
    hipLaunchKernelGGL((k_jacobi_patch<F, D, Cn>), dim3(pa.npatch), dim3(kPatch), sm, s, ta, pa.tloc, pa.pvoff, \
                       pa.pverts, pa.pslot, pa.inc_off, pa.inc, pos, pa.partial, cacc, nt, pa.maxl, (int)xcd)
    const int sel = (first ? 4 : 0) | (det ? 2 : 0) | (con ? 1 : 0);
    switch (sel) {
    case 0: XPBD_PATCH_LAUNCH(false, false, false); break;
    case 1: XPBD_PATCH_LAUNCH(false, false, true); break;
    case 2: XPBD_PATCH_LAUNCH(false, true, false); break;
    case 3: XPBD_PATCH_LAUNCH(false, true, true); break;
    case 4: XPBD_PATCH_LAUNCH(true, false, false); break;
    case 5: XPBD_PATCH_LAUNCH(true, false, true); break;
    case 6: XPBD_PATCH_LAUNCH(true, true, false); break;
    default: XPBD_PATCH_LAUNCH(true, true, true); break;
    }
#undef XPBD_PATCH_LAUNCH
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || pa.nshared <= 0) return e;
    if (con)
        hipLaunchKernelGGL(k_jacobi_shared<true>, dim3(nblk(pa.nshared, kBlock)), dim3(kBlock), 0, s, pa.shared,
                           pa.soff, pa.partial, cacc, pos, pa.nshared);
    else
        hipLaunchKernelGGL(k_jacobi_shared<false>, dim3(nblk(pa.nshared, kBlock)), dim3(kBlock), 0, s, pa.shared,
                           pa.soff, pa.partial, cacc, pos, pa.nshared);
    return hipGetLastError();
}

size_t jacobi_fused_smem(int maxl, bool con)
{
    return jacobi_patch_smem(maxl, true) + (con ? (size_t)maxl * 12 + (size_t)kPatch * 40 : 0);
}

hipError_t launch_jacobi_fused(hipStream_t s, const TetArrays& ta, const FusedPatchArgs& A, bool first, bool con)
{
    if (A.npatch <= 0) return hipSuccess;
    const size_t sm = jacobi_fused_smem(A.maxl, con);
    if (first && con) hipLaunchKernelGGL((k_jacobi_fused<true, true>), dim3(A.npatch), dim3(kPatch), sm, s, ta, A);
    else if (first) hipLaunchKernelGGL((k_jacobi_fused<true, false>), dim3(A.npatch), dim3(kPatch), sm, s, ta, A);
    else if (con) hipLaunchKernelGGL((k_jacobi_fused<false, true>), dim3(A.npatch), dim3(kPatch), sm, s, ta, A);
    else hipLaunchKernelGGL((k_jacobi_fused<false, false>), dim3(A.npatch), dim3(kPatch), sm, s, ta, A);
    return hipGetLastError();
}

hipError_t launch_jacobi_fused_final(hipStream_t s, const int* shared, const int* soff, const float4* prd,
                                     const float4* brd, float4* pos, int nshared)
{
    if (nshared <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_jacobi_fused_final, dim3(nblk(nshared, kBlock)), dim3(kBlock), 0, s, shared, soff, prd, brd,
                       pos, nshared);
    return hipGetLastError();
}

hipError_t launch_jacobi_apply(hipStream_t s, float4* pos, float* acc, int nv)
{
    if (nv <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_jacobi_apply, dim3(nblk(nv, kBlock)), dim3(kBlock), 0, s, pos,
                       reinterpret_cast<float4*>(acc), nv);
    return hipGetLastError();
}

hipError_t launch_fill_zero(hipStream_t s, void* p, size_t bytes)
{
    const size_t n = bytes / 4;
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_fill_zero, dim3(nblk((long long)n, kBlock)), dim3(kBlock), 0, s,
                       static_cast<uint32_t*>(p), n);
    return hipGetLastError();
}

hipError_t launch_residual(hipStream_t s, const TetArrays& ta, const float4* pos, const float4* lo, bool with_delta,
                           int nt, double* part, int nblocks, double* vec)
{
    hipLaunchKernelGGL(k_residual, dim3(nblocks), dim3(kBlock), 0, s, ta, pos, lo, with_delta ? 1 : 0, nt, part,
                       vec);
    return hipGetLastError();
}

hipError_t launch_energy(hipStream_t s, const TetArrays& ta, const float4* pos, const float4* lo, int nt, double dt2,
                         double* part, int nblocks)
{
    hipLaunchKernelGGL(k_energy, dim3(nblocks), dim3(kBlock), 0, s, ta, pos, lo, nt, dt2, part);
    return hipGetLastError();
}

hipError_t launch_sum_partials(hipStream_t s, const double* part, int n, double* out)
{
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kBlock), 0, s, part, n, out);
    return hipGetLastError();
}

hipError_t launch_keep_prediction(hipStream_t s, const float4* pos, float4* pred, int nv)
{
    if (nv <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_keep_prediction, dim3(nblk(nv, kBlock)), dim3(kBlock), 0, s, pos, pred, nv);
    return hipGetLastError();
}

hipError_t launch_element_forces(hipStream_t s, const TetArrays& ta, const PrimaryArgs& A)
{
    if (A.nt <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_element_forces, dim3(nblk(A.nt, kBlock)), dim3(kBlock), 0, s, ta, A);
    return hipGetLastError();
}

hipError_t launch_consistent_shift(hipStream_t s, float4* pos, const float4* pred, const int* vt_off,
                                   const double* ef, int nv)
{
    if (nv <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_consistent_shift, dim3(nblk(nv, kBlock)), dim3(kBlock), 0, s, pos, pred, vt_off, ef, nv);
    return hipGetLastError();
}

hipError_t launch_primary_residual(hipStream_t s, const TetArrays& ta, const PrimaryArgs& A, double* out)
{
    if (A.nv <= 0 || A.nblocks <= 0) return hipSuccess;
    if (A.nt > 0) hipLaunchKernelGGL(k_element_forces, dim3(nblk(A.nt, kBlock)), dim3(kBlock), 0, s, ta, A);
    hipLaunchKernelGGL(k_primary_residual, dim3(nblk(A.nv, kBlock)), dim3(kBlock), 0, s, ta, A);
    if (A.natt > 0)
        hipLaunchKernelGGL(k_zero_attached, dim3(nblk(A.natt, kBlock)), dim3(kBlock), 0, s, A.att_v, A.natt, A.nv,
                           A.R);
    hipLaunchKernelGGL(k_sum_squares, dim3(A.nblocks), dim3(kBlock), 0, s, A.R, 3 * (size_t)A.nv, A.part);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kBlock), 0, s, A.part, A.nblocks, out);
    return hipGetLastError();
}

hipError_t launch_detect(hipStream_t s, const int4* faces, int nf, const double* rest, const float4* pos,
                         const float4* lo, const SdfDev* sdf, const unsigned long long* step, double dt, int* flag,
                         double* cp, double* cn, double* cb)
{
    if (nf <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_detect, dim3(nblk(nf, kBlock)), dim3(kBlock), 0, s, faces, nf, rest, pos, lo, sdf, step, dt,
                       flag, cp, cn, cb);
    return hipGetLastError();
}

int compact_blocks(int nf) { return (int)nblk(nf > 0 ? nf : 1, kBlock); }

hipError_t launch_compact(hipStream_t s, const int* flag, int nf, int* cpos, int* clist, int* bsum)
{
    const int nb = compact_blocks(nf);
    hipLaunchKernelGGL(k_count_hits, dim3(nb), dim3(kBlock), 0, s, flag, nf, bsum);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(kScan), 0, s, bsum, nb, cpos, nf);
    hipLaunchKernelGGL(k_scatter_hits, dim3(nb), dim3(kBlock), 0, s, flag, nf, bsum, cpos, clist);
    return hipGetLastError();
}

hipError_t launch_contacts_gs(hipStream_t s, const int4* faces, const int* fci_off, int ncol, int ninst,
                              const int* cpos, const int* clist, const double* rest, float4* pos, const float4* lo,
                              const double* cp, const double* cn, const double* cb, double* clam, bool first,
                              float4* dsep, const float4* dbase)
{
    if (ninst <= 0 || ncol <= 0) return hipSuccess;
    const dim3 g(ninst), b(kBlock);
    const bool split = dsep && dbase;
    if (first && split)
        hipLaunchKernelGGL((k_contacts_gs<true, true>), g, b, 0, s, faces, fci_off, ncol, ninst, cpos, clist, rest, pos,
                           dsep, dbase, lo, cp, cn, cb, clam);
    else if (first)
        hipLaunchKernelGGL((k_contacts_gs<true, false>), g, b, 0, s, faces, fci_off, ncol, ninst, cpos, clist, rest,
                           pos, nullptr, nullptr, lo, cp, cn, cb, clam);
    else if (split)
        hipLaunchKernelGGL((k_contacts_gs<false, true>), g, b, 0, s, faces, fci_off, ncol, ninst, cpos, clist, rest,
                           pos, dsep, dbase, lo, cp, cn, cb, clam);
    else
        hipLaunchKernelGGL((k_contacts_gs<false, false>), g, b, 0, s, faces, fci_off, ncol, ninst, cpos, clist, rest,
                           pos, nullptr, nullptr, lo, cp, cn, cb, clam);
    return hipGetLastError();
}

hipError_t launch_contacts_jacobi(hipStream_t s, const int4* faces, int nf, const int* cpos, const int* clist,
                                  const double* rest, const float4* pos, const float4* lo, float* acc,
                                  const double* cp, const double* cn, const double* cb, double* clam, bool first)
{
    if (nf <= 0) return hipSuccess;
    const dim3 g(nblk(nf, kBlock));
    if (first)
        hipLaunchKernelGGL(k_contacts_jacobi<true>, g, dim3(kBlock), 0, s, faces, nf, cpos, clist, rest, pos, lo,
                           acc, cp, cn, cb, clam);
    else
        hipLaunchKernelGGL(k_contacts_jacobi<false>, g, dim3(kBlock), 0, s, faces, nf, cpos, clist, rest, pos, lo,
                           acc, cp, cn, cb, clam);
    return hipGetLastError();
}

}  // namespace xpbd
