// gfx950 kernels of the eqdyna_amd explicit time step.
//
// One step of the reference (src/driver.f90:9-34) becomes
//   station sample -> element kernels -> fault kernel -> node kernels
// where the node kernels fuse "F/m" of this step with the central-difference
// update of the next one (velDispUpdate, src/driver.f90:89-155), so the
// assembled force never round-trips through HBM.  The summation order inside
// every element and every node follows the reference; the only regrouping is
// that an element's stiffness and hourglass forces on a node are combined
// before the node sums its elements (the reference sums all element
// stiffness forces first, then all hourglass forces).
//
// Compiled with -ffp-contract=off: no FMA contraction, so each rounded
// operation is the one the reference performs.
#include <hip/hip_ext.h>

#include "eqd_device.h"

namespace eqd {

template <typename R>
__device__ __forceinline__ R sqr(R x) { return x * x; }


// Element forces and stresses are stored non-temporal: a force is read once,
// by the node kernels, a stress by its own element a step later, and
// bypassing the caches keeps the node state (gathered by the next step's
// element kernels) resident in the Infinity Cache instead.  Measured on the
// rough-fault box (forces): FP32 step 0.718 -> 0.690 ms, FP64 3.7 % faster,
// node kernel 0.159 -> 0.139 ms; the stresses add 0.2-0.6 %.
template <typename R>
__device__ __forceinline__ void store_nt(R *p, R v)
{
    __builtin_nontemporal_store(v, p);
}

// two values of a force slot or summed-stress row (pml_at, reg_ef_at) as one access in
// FP64, as two in FP32
template <typename R>
struct Pair;
template <>
struct Pair<double> {
    typedef double T __attribute__((ext_vector_type(2)));
};

template <typename R>
__device__ __forceinline__ void store_pair_nt(R *p, R a, R b)
{
    if constexpr (sizeof(R) == sizeof(double)) {
        const typename Pair<R>::T v = {a, b};
        __builtin_nontemporal_store(v, (typename Pair<R>::T *)p);
    } else {
        __builtin_nontemporal_store(a, p);
        __builtin_nontemporal_store(b, p + 64);
    }
}

template <typename R>
__device__ __forceinline__ void store_pair(R *p, R a, R b)
{
    if constexpr (sizeof(R) == sizeof(double)) {
        const typename Pair<R>::T v = {a, b};
        *(typename Pair<R>::T *)p = v;
    } else {
        p[0] = a;
        p[64] = b;
    }
}

template <typename R>
__device__ __forceinline__ void load_pair(const R *p, R &a, R &b)
{
    if constexpr (sizeof(R) == sizeof(double)) {
        const typename Pair<R>::T v = *(const typename Pair<R>::T *)p;
        a = v.x;
        b = v.y;
    } else {
        a = p[0];
        b = p[64];
    }
}

// The element kernels' other once-per-step streams (IEN, stresses, operator
// rows) are loaded non-temporal for the same reason: c4 FP32 step 0.688 ->
// 0.664 ms, FP64 3.2 % faster, c2 FP32 1.1 %.
template <typename T>
__device__ __forceinline__ T load_nt(const T *p)
{
    return __builtin_nontemporal_load(p);
}

// Node state in the precision of the mode (R): FP64 arrays, or FP32
// velocity and the FP32 displacement pair (eqd_device.h, NodeState).
template <typename R>
__device__ __forceinline__ double node_v(const NodeState &N, int node, int k)
{
    if constexpr (sizeof(R) == sizeof(float)) return (double)N.vel32[blk_at<3>(node, k)];
    else return N.vel[blk_at<3>(node, k)];
}

// bfloat16 bits <-> float (round to nearest even; the residuals converted
// here are finite)
__device__ __forceinline__ float bf16_float(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ uint16_t float_bf16(float x)
{
    const uint32_t u = __float_as_uint(x);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

template <typename R>
__device__ __forceinline__ double node_u(const NodeState &N, int node, int k)
{
    if constexpr (sizeof(R) == sizeof(float))
        return (double)N.dhi[blk_at<3>(node, k)] + (double)bf16_float(N.dlo[blk_at<3>(node, k)]);
    else return N.disp[blk_at<3>(node, k)];
}

// a PML node's damping coefficient: FP64, or in the FP32 bulk mode its FP32
// rounding (it scales the FP32 split velocities)
template <typename R>
__device__ __forceinline__ double pml_damp(const PmlNodes &P, int64_t q, int k)
{
    if constexpr (sizeof(R) == sizeof(float)) return (double)P.damp32[blk_at<3>(q, k)];
    else return P.damp[blk_at<3>(q, k)];
}

// the nodal mass as the node kernels divide by it: FP64, or in the FP32 bulk
// mode its FP32 rounding (a relative change of at most 2^-24, below the
// rounding of the FP32 velocity it feeds)
template <typename R>
__device__ __forceinline__ double node_mass(const NodeState &N, int node)
{
    if constexpr (sizeof(R) == sizeof(float)) return (double)N.mass32[node];
    else return N.mass[node];
}

// velocity and displacement of node component k as the element kernels
// gather them: the stored values in the mode's precision (FP32 bulk: the
// rounded displacement dhi)
template <typename R>
__device__ __forceinline__ R gather_v(const NodeState &N, int node, int k)
{
    if constexpr (sizeof(R) == sizeof(float)) return N.vel32[blk_at<3>(node, k)];
    else return N.vel[blk_at<3>(node, k)];
}

template <typename R>
__device__ __forceinline__ R gather_u(const NodeState &N, int node, int k)
{
    if constexpr (sizeof(R) == sizeof(float)) return N.dhi[blk_at<3>(node, k)];
    else return N.disp[blk_at<3>(node, k)];
}

// store a node component's new velocity and displacement (computed in FP64)
template <typename R>
__device__ __forceinline__ void node_store(const NodeState &N, int node, int k, double v, double u)
{
    const int64_t j = blk_at<3>(node, k);
    if constexpr (sizeof(R) == sizeof(float)) {
        const float hi = (float)u;
        N.vel32[j] = (float)v;
        N.dhi[j] = hi;
        N.dlo[j] = float_bf16((float)(u - (double)hi));
    } else {
        N.vel[j] = v;
        N.disp[j] = u;
    }
}

// ------------------------------------------------------------------------
// Operators of one hexahedron rebuilt from its corner coordinates (kOpsGeom):
// the quantities of the one-time precompute (calcGlobalShapeFunc, vlm and
// calcSSPhi4Hrgls; src/calcGlobalShapeFunc.f90:1-75, src/library.f90:60-93,
// src/assembleGlobalMass.f90:328-374) in the kernel's precision, from the
// coordinates of local nodes 1..7 relative to node 0, so the values are
// rounded relative to the element's size.  dN/dx at the centre, the Jacobian
// determinant, the hourglass stiffness SS and the compact hourglass
// operators (1/s_m, h_m/s_m; phi_compact).
template <typename R>
struct GeoOps {
    R shp[24];
    R det;
    R ss[6];
    R cph[16];
};

template <typename R>
__device__ __forceinline__ void geo_ops(const R *opb, int64_t t, R lam, R mu, GeoOps<R> &G)
{
    // sign of the trilinear shape function of node a in direction i at the centre
    constexpr int ac[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                              {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
    R x[3][8];
#pragma unroll
    for (int j = 0; j < 3; j++) x[j][0] = 0;
#pragma unroll
    for (int a = 1; a < 8; a++)
#pragma unroll
        for (int j = 0; j < 3; j++) x[j][a] = opb_at<R, kOpG>(opb, t, (a - 1) * 3 + j);
    // xs[j][i] = d x_j / d xi_i = sum_a ac[a][i] x_j(a) / 8
    R xs[3][3];
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
        for (int i = 0; i < 3; i++) {
            R s = 0;
#pragma unroll
            for (int a = 1; a < 8; a++) s = ac[a][i] > 0 ? s + x[j][a] : s - x[j][a];
            xs[j][i] = s * (R)0.125;
        }
    const R c11 = xs[1][1] * xs[2][2] - xs[1][2] * xs[2][1];
    const R c12 = xs[1][2] * xs[2][0] - xs[1][0] * xs[2][2];
    const R c13 = xs[1][0] * xs[2][1] - xs[1][1] * xs[2][0];
    const R c21 = xs[2][1] * xs[0][2] - xs[2][2] * xs[0][1];
    const R c22 = xs[2][2] * xs[0][0] - xs[2][0] * xs[0][2];
    const R c23 = xs[2][0] * xs[0][1] - xs[2][1] * xs[0][0];
    const R c31 = xs[0][1] * xs[1][2] - xs[0][2] * xs[1][1];
    const R c32 = xs[0][2] * xs[1][0] - xs[0][0] * xs[1][2];
    const R c33 = xs[0][0] * xs[1][1] - xs[0][1] * xs[1][0];
    const R det = xs[0][0] * c11 + xs[0][1] * c12 + xs[0][2] * c13;
    const R rdet = (R)1 / det;
    const R q = (R)0.125 * rdet;
#pragma unroll
    for (int a = 0; a < 8; a++) {
        const R s1 = (R)ac[a][0], s2 = (R)ac[a][1], s3 = (R)ac[a][2];
        G.shp[a * 3 + 0] = (s1 * c11 + s2 * c12 + s3 * c13) * q;
        G.shp[a * 3 + 1] = (s1 * c21 + s2 * c22 + s3 * c23) * q;
        G.shp[a * 3 + 2] = (s1 * c31 + s2 * c32 + s3 * c33) * q;
    }
    G.det = det;
    // element volume (vlm, src/library.f90:60-93)
    constexpr int it[8][8] = {{1, 2, 3, 4, 5, 6, 7, 8}, {2, 3, 4, 1, 6, 7, 8, 5}, {3, 4, 1, 2, 7, 8, 5, 6},
                              {4, 1, 2, 3, 8, 5, 6, 7}, {5, 8, 7, 6, 1, 4, 3, 2}, {6, 5, 8, 7, 2, 1, 4, 3},
                              {7, 6, 5, 8, 3, 2, 1, 4}, {8, 7, 6, 5, 4, 3, 2, 1}};
    R vol = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
#define GY(r) x[1][it[i][(r) - 1] - 1]
#define GZ(r) x[2][it[i][(r) - 1] - 1]
        const R bb = GY(2) * (GZ(6) - GZ(3) + GZ(5) - GZ(4)) + GY(3) * (GZ(2) - GZ(4)) +
                     GY(4) * (GZ(3) - GZ(8) + GZ(2) - GZ(5)) + GY(5) * (GZ(8) - GZ(6) + GZ(4) - GZ(2)) +
                     GY(6) * (GZ(5) - GZ(2)) + GY(8) * (GZ(4) - GZ(5));
#undef GY
#undef GZ
        vol = vol + x[0][i] * bb;
    }
    vol = vol / (R)12;
    // SS = co * X^T X with X the inverse Jacobian (calcSSPhi4Hrgls)
    const R ce = (R)16 * (mu * ((R)3 * lam + (R)2 * mu) / (lam + mu)) / (R)15;
    const R co = ce * vol / (R)48;
    const R X[3][3] = {{c11 * rdet, c21 * rdet, c31 * rdet}, {c12 * rdet, c22 * rdet, c32 * rdet},
                       {c13 * rdet, c23 * rdet, c33 * rdet}};
    G.ss[0] = co * (X[0][0] * X[0][0] + X[1][0] * X[1][0] + X[2][0] * X[2][0]);
    G.ss[1] = co * (X[0][0] * X[0][1] + X[1][0] * X[1][1] + X[2][0] * X[2][1]);
    G.ss[2] = co * (X[0][0] * X[0][2] + X[1][0] * X[1][2] + X[2][0] * X[2][2]);
    G.ss[3] = co * (X[0][1] * X[0][1] + X[1][1] * X[1][1] + X[2][1] * X[2][1]);
    G.ss[4] = co * (X[0][1] * X[0][2] + X[1][1] * X[1][2] + X[2][1] * X[2][2]);
    G.ss[5] = co * (X[0][2] * X[0][2] + X[1][2] * X[1][2] + X[2][2] * X[2][2]);
    // compact hourglass operators: h_m = sum_b Gamma_mb x_b, s_m the rms over
    // a of Gamma_ma - h_m . dN_a/dx
#pragma unroll
    for (int m = 0; m < 4; m++) {
        R h[3];
#pragma unroll
        for (int j = 0; j < 3; j++) {
            R s = 0;
#pragma unroll
            for (int b = 1; b < 8; b++) s = ((kGammaNeg[m] >> b) & 1u) ? s - x[j][b] : s + x[j][b];
            h[j] = s;
        }
        R sq = 0;
#pragma unroll
        for (int a = 0; a < 8; a++) {
            const R g = ((kGammaNeg[m] >> a) & 1u) ? (R)-1 : (R)1;
            const R ph = g - (h[0] * G.shp[a * 3] + h[1] * G.shp[a * 3 + 1] + h[2] * G.shp[a * 3 + 2]);
            sq = sq + ph * ph;
        }
        const R rs = (R)1 / sqrt(sq / (R)8);
        G.cph[m] = rs;
#pragma unroll
        for (int j = 0; j < 3; j++) G.cph[4 + 3 * m + j] = h[j] * rs;
    }
}

// split velocity k of PML node q
template <typename R>
__device__ __forceinline__ double pml_v(const PmlNodes &P, int64_t q, int k)
{
    if constexpr (sizeof(R) == sizeof(float)) return (double)P.v12f[blk_at<12>(q, k)];
    else return P.v12[blk_at<12>(q, k)];
}

template <typename R>
__device__ __forceinline__ void pml_store(const PmlNodes &P, int64_t q, int k, double v)
{
    if constexpr (sizeof(R) == sizeof(float)) P.v12f[blk_at<12>(q, k)] = (float)v;
    else P.v12[blk_at<12>(q, k)] = v;
}

// A node's state at full precision: the FP32 bulk mode keeps split nodes in
// the fault's FP64 arrays (pair_slot says where); every other node is read
// as node_v / node_u read it.
template <typename R>
__device__ __forceinline__ double exact_v(const Member<R> &M, int node, int k)
{
    if constexpr (sizeof(R) == sizeof(float)) {
        const int ps = M.N.pair_slot[node];
        if (ps >= 0) return M.F.sv[((ps & 1) * 3 + k) * M.F.ld + (ps >> 1)];
    }
    return node_v<R>(M.N, node, k);
}

template <typename R>
__device__ __forceinline__ double exact_u(const Member<R> &M, int node, int k)
{
    if constexpr (sizeof(R) == sizeof(float)) {
        const int ps = M.N.pair_slot[node];
        if (ps >= 0) return M.F.su[((ps & 1) * 3 + k) * M.F.ld + (ps >> 1)];
    }
    return node_u<R>(M.N, node, k);
}

// The start of a step for one scenario, by one block: the clock advance of
// src/driver.f90:9-11 (nt += 1, timeElapsed += dt with the reference's FP64
// addition, and the run's last-step flag), then storeOffFaultStData
// (src/driver.f90:157-180) at the new clock: time and the recorded
// displacement or velocity of every off-fault station row.  Every thread
// reads the old clock before the barrier and thread 0 writes the new one
// after it.  Later kernels of the step read the advanced clock from device
// memory; the element kernels do not read it.
template <typename R>
__device__ __forceinline__ void step_start_block(const Member<R> &M)
{
    Clock *c = M.clk;
    const int nt = c->nt + 1;
    const double time = c->time + M.C.dt;
    const int run_end = c->run_end;
    __syncthreads();
    if (threadIdx.x == 0) {
        c->nt = nt;
        c->time = time;
        c->mode = nt == run_end ? 1 : 0;
    }
    const OffStations S = M.S;
    if (S.nrow <= 0) return;
    double *col = S.hist + (int64_t)(nt - 1) * (S.nrow + 1);
    for (int r = threadIdx.x; r <= S.nrow; r += blockDim.x) {
        if (r == 0) {
            col[0] = time;
            continue;
        }
        const int i = r - 1;
        const int node = S.idh[i], k = S.idh[S.nrow + i], kind = S.idh[2 * S.nrow + i];
        if (kind == 1)
            col[r] = exact_u(M, node, k);
        else if (kind == 2)
            col[r] = exact_v(M, node, k);
    }
}

// store a node's new state at full precision (and its rounded copy)
template <typename R>
__device__ __forceinline__ void exact_store(const Member<R> &M, int node, int k, double v, double u)
{
    if constexpr (sizeof(R) == sizeof(float)) {
        const int ps = M.N.pair_slot[node];
        if (ps >= 0) {
            const int64_t j = ((ps & 1) * 3 + k) * M.F.ld + (ps >> 1);
            M.F.sv[j] = v;
            M.F.su[j] = u;
        }
    }
    node_store<R>(M.N, node, k, v, u);
}

// ------------------------------------------------------------------------
// Regular (type 1, 11, 12, 13) elements: calcElemKU + calcElemMass + hrglss
// for one element per thread.
//   src/assembleGlobalKU.f90:13-35, src/calcElemKU.f90:3-191,
//   src/calcElemMass.f90:1-19, src/hrglss.f90:13-98
// kQ: C_Q == 1; kDP: C_elastic == 0; kHG: C_hg (1 stiffness, 2 viscous);
// kMass: inertia/gravity term present (rdampm != 0 or plastic).
//
// Streaming form: one pass over the eight nodes gathers each node's v and u
// once and accumulates everything that needs them (strain rate, strain, the
// four hourglass mode amplitudes), so no per-node array stays live; the
// stress update follows; a second pass over the nodes forms and stores each
// node's force.  Every accumulation keeps the reference's order (nodes in
// ascending order, modes in ascending order), and the register footprint is
// a few dozen values instead of four 24-vectors.
// Nodes whose v and u are in flight ahead of the arithmetic in the element
// kernels' gather pass: more hides more memory latency per wave, fewer keeps
// the register footprint (and so the number of waves per SIMD) down.
#ifndef EQD_PREFETCH
#define EQD_PREFETCH 1
#endif
// Unrolling of the element kernels' force pass.
#ifndef EQD_ELEM_UNROLL
#define EQD_ELEM_UNROLL 1
#endif
// Operators read from memory (no LDS class tables): keep each node's operator
// values in registers from the gather pass to the force pass (both passes
// unrolled), so they are read from memory once per step instead of twice --
// 2 in both precisions, 1 in FP32 only, 0 never (the force pass reloads them).
// Measured on the rough-fault box: regular elements 10 % (FP64) and 15 %
// (FP32) faster when kept; the PML kernel, whose 12-slot force pass needs
// more registers, slows down when kept (one wave per SIMD), so it reloads.
#ifndef EQD_KEEP_OPS
#define EQD_KEEP_OPS 2
#endif
#ifndef EQD_KEEP_OPS_PML
#define EQD_KEEP_OPS_PML 0
#endif
// PML kernel with operators from memory: 1 stages each thread's dN/dx (24
// values) in LDS between the passes (its own slots, no barrier), so the force
// pass re-reads only Phi from memory.
#ifndef EQD_STAGE_SHP_PML
#define EQD_STAGE_SHP_PML 1
#endif
// Waves per SIMD the geometry-row variants (kOpsGeom) are bounded to.
// Measured on the rough-fault box (FP32, 600 steps): regular elements
// 0.150 ms unbounded (130 VGPRs, 3 waves) against 0.138 ms at 4 waves; PML
// elements 0.154 ms at 4 waves (spilling) against 0.115 ms at 3 waves.  The
// regular variants with Q, plasticity or an inertia term would spill at 4
// waves and stay unbounded.
#ifndef EQD_REG_GEO_WAVES
#define EQD_REG_GEO_WAVES 4
#endif
#ifndef EQD_PML_GEO_WAVES
#define EQD_PML_GEO_WAVES 3
#endif
// Waves per SIMD the fault kernel is bounded to (1: unbounded, 256 VGPRs).
// Bounded to 2 (28-68 bytes per lane spilled) it measured the same on the
// 64-scenario ensemble (12.90 against 12.88 ms per step, ~9,650 fault waves)
// and on the rough-fault box (0.552 against 0.550 ms): the pairs' kernel runs
// beside the node kernels and fills the slots they leave, so its occupancy is
// not what bounds the step.
#ifndef EQD_FAULT_WAVES
#define EQD_FAULT_WAVES 1
#endif

//
// kOps: where the operators come from (kOpsRows, kOpsLds, kOpsCompact).  With
// kOpsLds both class tables are staged in LDS (dynamic shared memory of
// (n_opc * kOpW + n_phic * 32) values) at block start, so the operator reads
// of the element loop are LDS reads; with kOpsCompact Phi is rebuilt per node
// (phi_compact) from 16 values read once per element; with kOpsGeom all
// operators are rebuilt from the element's corner coordinates (geo_ops).
template <typename R, bool kQ, bool kDP, int kHG, bool kMass, int kOps>
__global__ void __launch_bounds__(256, (kOps == kOpsGeom && !kQ && !kDP && !kMass ? EQD_REG_GEO_WAVES : 1))
    elem_regular_kernel(const Member<R> *__restrict__ Ms, const uint16_t *__restrict__ bmem, int remap, int nstart)
{
    // With nstart > 0 the launch carries the step start of scenarios
    // 0 .. nstart-1 in its first blocks, rounded up to a multiple of 8 so
    // that the element blocks keep their XCD assignment.
    const int nskip = step_start_skip(nstart);
    if ((int)blockIdx.x < nskip) {
        if ((int)blockIdx.x < nstart) step_start_block<R>(Ms[blockIdx.x]);
        return;
    }
    constexpr bool kGeo = kOps == kOpsGeom;
    constexpr bool kLds = kOps == kOpsLds, kCmp = kOps == kOpsCompact || kGeo;
    constexpr int kW = kGeo ? kOpG : (kCmp ? kOpBc : kOpB);
    constexpr int kNst = kQ ? 12 : 6;  // stress values per element (with the Q memory variables)
    extern __shared__ __attribute__((aligned(16))) unsigned char eqd_smem[];
    typedef __attribute__((address_space(3))) R lds_R;
    const unsigned gb = launch_block(remap, (unsigned)nskip);
    const Member<R> &M = Ms[member_of(bmem, gb)];
    const RegElems<R> E = M.E;
    const NodeState N = M.N;
    const StepConsts C = M.C;
    const int64_t t = (int64_t)(gb - (unsigned)M.blk0[kKElem]) * blockDim.x + threadIdx.x;
    const R *op = nullptr, *pht = nullptr;
    const lds_R *lrow = nullptr, *lphr = nullptr;
    if constexpr (kLds) {
        lds_R *lop = (lds_R *)eqd_smem, *lph = lop + E.n_opc * kOpW;
        for (int j = threadIdx.x; j < E.n_opc * kOpW; j += blockDim.x) lop[j] = E.optab[j];
        for (int j = threadIdx.x; j < E.n_phic * 32; j += blockDim.x) lph[j] = E.phitab[j];
        __syncthreads();
        if (t >= E.n) return;
        lrow = lop + (int64_t)E.opc[t] * kOpW;
        lphr = lph + (int64_t)E.phic[t] * 32;
    } else {
        if (t >= E.n) return;
        // operators: one class-table row, or the element's own arrays
        op = E.optab ? E.optab + (int64_t)E.opc[t] * kOpW : nullptr;
        pht = E.phitab ? E.phitab + (int64_t)E.phic[t] * 32 : nullptr;
    }
    const int64_t n = E.n, nn = N.nn;
    constexpr bool kF32 = sizeof(R) != sizeof(double);
    GeoOps<R> geo;
    if constexpr (kGeo)
        geo_ops<R>(E.opb, t, opb_at<R, kOpG>(E.opb, t, kGeoMat), opb_at<R, kOpG>(E.opb, t, kGeoMat + 1), geo);
    auto SHP = [&](int i) -> R {
        if constexpr (kGeo) return geo.shp[i];
        else if constexpr (kLds) return lrow[kOpShp + i];
        else if constexpr (kCmp) return opb_at<R, kW>(E.opb, t, kOpShp + i);
        else return op ? op[kOpShp + i] : opb_at(E.opb, t, kOpShp + i);
    };
    auto PHI = [&](int m, int a) -> R {
        if constexpr (kLds) return lphr[m * 8 + a];
        else return pht ? pht[m * 8 + a] : opb_at(E.opb, t, kOpW + m * 8 + a);
    };
    // scalar operator k of the element's row
    auto OPK = [&](int k) -> R {
        if constexpr (kGeo) {
            if (k == kOpDet) return geo.det;
            if (k >= kOpSs) return geo.ss[k - kOpSs];
            return opb_at<R, kOpG>(E.opb, t, kGeoMat + (k - kOpLam));
        } else if constexpr (kLds) {
            return lrow[k];
        } else if constexpr (kCmp) {
            return opb_at<R, kW>(E.opb, t, k);
        } else {
            return op ? op[k] : opb_at(E.opb, t, k);
        }
    };
    // compact hourglass operators (kOpsCompact or kOpsGeom with the stiffness hourglass)
    constexpr bool kCHg = kCmp && kHG == 1;
    R cph[kCHg ? 16 : 1];
    if constexpr (kCHg) {
#pragma unroll
        for (int i = 0; i < 16; i++) {
            if constexpr (kGeo) cph[i] = geo.cph[i];
            else cph[i] = opb_at<R, kW>(E.opb, t, kOpW + i);
        }
    }
    const R rdk = (R)C.rdampk;
    // all eight node ids first (independent loads), then the nodes' v and u
    // one node ahead of the arithmetic, so a gather is always in flight
    int nds[8];
#pragma unroll
    for (int a = 0; a < 8; a++) nds[a] = load_nt(E.ien + blk_at<8>(t, a));
    // EQD_PREFETCH nodes in flight: slot a % EQD_PREFETCH holds node a's v and u
    constexpr int P = EQD_PREFETCH;
    R vq[P][3], uq[P][3];
#pragma unroll
    for (int b = 0; b < P; b++)
#pragma unroll
        for (int k = 0; k < 3; k++) {
            vq[b][k] = gather_v<R>(N, nds[b], k);
            uq[b][k] = gather_u<R>(N, nds[b], k);
        }
    // FP32 bulk: velocities and displacements relative to the first node.
    // Strain, strain rate and the hourglass modes are blind to a common
    // translation (sum_a dN_a/dx = 0, Phi orthogonal to constants), so only
    // the rounding changes: the arithmetic then rounds relative to the
    // deformation rather than to the absolute motion.
    R v0[3] = {0, 0, 0}, u0[3] = {0, 0, 0};
    if constexpr (kF32) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            v0[k] = vq[0][k];
            u0[k] = uq[0][k];
        }
    }
    // pass 1: strain rate, strain (Q), hourglass mode amplitudes
    R sr[6] = {0, 0, 0, 0, 0, 0}, strain[6] = {0, 0, 0, 0, 0, 0};
    R pd[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};  // kHG 1: sum_a Phi_ma (u + rdampk v)_a
    R qv[3][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}};    // kHG 2: sum_a v_a Gamma_ja
    constexpr int fi[4][8] = {{1, 1, -1, -1, -1, -1, 1, 1}, {1, -1, -1, 1, -1, 1, 1, -1},
                              {1, -1, 1, -1, 1, -1, 1, -1}, {1, -1, 1, -1, -1, 1, -1, 1}};
    // a node's operator values (dN/dx, and Phi of the four modes), loaded one
    // node ahead like its v and u
    constexpr int kNo = kHG == 1 ? 7 : 3;
    auto set_phi = [&](int a, R o[kNo]) {
        if constexpr (kCHg) {
#pragma unroll
            for (int m = 0; m < 4; m++) o[3 + m] = phi_compact<R>(cph, m, a, o[0], o[1], o[2]);
        }
    };
    auto load_op = [&](int a, R o[kNo]) {
#pragma unroll
        for (int k = 0; k < 3; k++) o[k] = SHP(a * 3 + k);
        if constexpr (kHG == 1 && !kCmp) {
#pragma unroll
            for (int m = 0; m < 4; m++) o[3 + m] = PHI(m, a);
        }
    };
    // (from LDS the values are read where they are used; kept, only dN/dx
    // when Phi is rebuilt from the compact operators)
    constexpr bool kKeep = !kLds && !kGeo && (EQD_KEEP_OPS == 2 || (EQD_KEEP_OPS == 1 && sizeof(R) == sizeof(float)));
    constexpr int kNk = kCmp ? 3 : kNo;
    R opc_[kNo], opn_[kNo];
    R keep[kKeep ? 8 : 1][kNk];
    if constexpr (!kLds) load_op(0, opc_);
constexpr int kU1 = (kKeep || kGeo) ? 8 : EQD_PREFETCH;
#pragma unroll kU1
    for (int a = 0; a < 8; a++) {
        const int b = a % P;
        if constexpr (kKeep) {
#pragma unroll
            for (int k = 0; k < kNk; k++) keep[a][k] = opc_[k];
        }
        set_phi(a, opc_);
        R vl[3], dl[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            vl[k] = vq[b][k] - v0[k];
            dl[k] = uq[b][k] - u0[k];
        }
        if (a + P < 8) {
            const int nx = nds[a + P];
#pragma unroll
            for (int k = 0; k < 3; k++) {
                vq[b][k] = gather_v<R>(N, nx, k);
                uq[b][k] = gather_u<R>(N, nx, k);
            }
        }
        if constexpr (kLds) {
            load_op(a, opc_);
        } else {
            if (a < 7) load_op(a + 1, opn_);
        }
        const R n1 = opc_[0], n2 = opc_[1], n3 = opc_[2];
        sr[0] = sr[0] + n1 * vl[0];
        sr[1] = sr[1] + n2 * vl[1];
        sr[2] = sr[2] + n3 * vl[2];
        sr[3] = sr[3] + n3 * vl[1] + n2 * vl[2];
        sr[4] = sr[4] + n3 * vl[0] + n1 * vl[2];
        sr[5] = sr[5] + n2 * vl[0] + n1 * vl[1];
        if (kQ) {
            strain[0] = strain[0] + n1 * dl[0];
            strain[1] = strain[1] + n2 * dl[1];
            strain[2] = strain[2] + n3 * dl[2];
            strain[3] = strain[3] + n3 * dl[1] + n2 * dl[2];
            strain[4] = strain[4] + n3 * dl[0] + n1 * dl[2];
            strain[5] = strain[5] + n2 * dl[0] + n1 * dl[1];
        }
        if (kHG == 1) {
            R ut[3];
#pragma unroll
            for (int k = 0; k < 3; k++) ut[k] = dl[k] + rdk * vl[k];
#pragma unroll
            for (int m = 0; m < 4; m++) {
                const R ph = opc_[3 + m];
#pragma unroll
                for (int k = 0; k < 3; k++) pd[m][k] = pd[m][k] + ph * ut[k];
            }
        } else if (kHG == 2) {
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) qv[i][j] = qv[i][j] + vl[i] * (R)fi[j][a];
        }
        if constexpr (!kLds) {
#pragma unroll
            for (int k = 0; k < kNo; k++) opc_[k] = opn_[k];
        }
    }
    const R lam = OPK(kOpLam), miu = OPK(kOpMu);
    const R det = OPK(kOpDet);
    const R cdiag = lam + (R)2 * miu;
    R srt[6];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        R acc = 0;
#pragma unroll
        for (int j = 0; j < 3; j++) acc = acc + (i == j ? cdiag : lam) * sr[j];
        srt[i] = acc;
    }
#pragma unroll
    for (int i = 3; i < 6; i++) srt[i] = miu * sr[i];

    R stress[6], strdev[6];
#pragma unroll
    for (int i = 0; i < 6; i++) stress[i] = load_nt(E.stress + blk_at<kNst>(t, i));
    if (!kQ) {
#pragma unroll
        for (int i = 0; i < 6; i++) {
            stress[i] = stress[i] + srt[i] * (R)C.dt;
            strdev[i] = stress[i];
        }
    } else {
        const double *q = E.qtab + 6 * E.qidx[t];
        const R e1 = (R)q[0], e2 = (R)q[1], wkp = (R)q[2], wks = (R)q[3], cv = (R)q[4], cs = (R)q[5];
        const R miuu = miu * cs;
        const R Mu = (lam + (R)2 * miu) * cv;
        const R vols = strain[0] + strain[1] + strain[2];
        R anestr1[6], anestr[6];
#pragma unroll
        for (int i = 0; i < 6; i++) anestr1[i] = load_nt(E.stress + blk_at<kNst>(t, 6 + i));
#pragma unroll
        for (int i = 0; i < 3; i++)
            anestr[i] = e1 * anestr1[i] + e2 * ((R)2 * miuu * strain[i] * wks + (Mu * wkp - (R)2 * miuu * wks) * vols);
#pragma unroll
        for (int i = 3; i < 6; i++) anestr[i] = e1 * anestr1[i] + e2 * (miuu * strain[i] * wks);
#pragma unroll
        for (int i = 0; i < 6; i++) store_nt(E.stress + blk_at<kNst>(t, 6 + i), anestr[i]);
#pragma unroll
        for (int i = 0; i < 3; i++)
            stress[i] = (R)2 * miuu * strain[i] + (Mu - (R)2 * miuu) * vols - (R)0.5 * (anestr[i] + anestr1[i]);
#pragma unroll
        for (int i = 3; i < 6; i++)
            stress[i] = (R)2 * miuu * strain[i] / (R)2 - (R)0.5 * (anestr[i] + anestr1[i]);
#pragma unroll
        for (int i = 0; i < 6; i++) strdev[i] = stress[i];
    }
    if (kDP) {
        // Drucker-Prager viscoplastic return, src/calcElemKU.f90:133-167
        const R strmea = (stress[0] + stress[1] + stress[2]) / (R)3;
#pragma unroll
        for (int i = 0; i < 3; i++) strdev[i] = stress[i] - strmea;
        R taomax = (R)0.5 * (sqr(strdev[0]) + sqr(strdev[1]) + sqr(strdev[2])) + sqr(strdev[3]) +
                   sqr(strdev[4]) + sqr(strdev[5]);
        taomax = sqrt(taomax);
        R yield = (R)C.ccosphi - (R)C.sinphi * (strmea + (R)E.porep[t]);
        if (yield < (R)0) yield = 0;
        R pstrmag = 0;
        if (taomax > yield) {
            const R rjust = yield / taomax + ((R)1 - yield / taomax) * (R)C.exp_dt_tv;
            R pstrinc[6];
#pragma unroll
            for (int i = 0; i < 6; i++) {
                stress[i] = strdev[i] * rjust;
                pstrinc[i] = (strdev[i] - stress[i]) / miu;
                if (i < 3) stress[i] = stress[i] + strmea;
            }
            const R pstrmea = (pstrinc[0] + pstrinc[1] + pstrinc[2]) / (R)3;
#pragma unroll
            for (int i = 0; i < 6; i++) pstrinc[i] = pstrinc[i] - pstrmea;
            R pm = (R)0.5 * (sqr(pstrinc[0]) + sqr(pstrinc[1]) + sqr(pstrinc[2])) + sqr(pstrinc[3]) +
                   sqr(pstrinc[4]) + sqr(pstrinc[5]);
            pstrmag = sqrt(pm);
        }
        E.pstrain[t] = E.pstrain[t] + (double)pstrmag;
    }
#pragma unroll
    for (int i = 0; i < 6; i++) store_nt(E.stress + blk_at<kNst>(t, i), stress[i]);

    // f = -m*al (calcElemMass), + (-det*w) * B^T (sigma + rdampk*sigma_dot), - hourglass
    const R temp = -det * (R)C.w;
    R st[6];
#pragma unroll
    for (int i = 0; i < 6; i++) st[i] = temp * (stress[i] + rdk * srt[i]);
    R g[4][3];  // kHG 1: SS * mode amplitude
    R coef = 0;
    if (kHG == 1) {
        R ss[6];
#pragma unroll
        for (int i = 0; i < 6; i++) ss[i] = OPK(kOpSs + i);
#pragma unroll
        for (int m = 0; m < 4; m++) {
            g[m][0] = ss[0] * pd[m][0] + ss[1] * pd[m][1] + ss[2] * pd[m][2];
            g[m][1] = ss[1] * pd[m][0] + ss[3] * pd[m][1] + ss[4] * pd[m][2];
            g[m][2] = ss[2] * pd[m][0] + ss[4] * pd[m][1] + ss[5] * pd[m][2];
        }
    } else if (kHG == 2) {
        const R rho = OPK(kOpRho), vp = OPK(kOpVp);
        coef = (R)0.25 * (R)C.kapa_hg * rho * vp * (R)pow((double)det * C.w, 2.0 / 3.0);
    }
    // pass 2: one node's force at a time (operators again one node ahead, or
    // kept from pass 1)
    if constexpr (!kLds && !kKeep) load_op(0, opc_);
constexpr int kU2 = (kKeep || kGeo) ? 8 : EQD_ELEM_UNROLL;
#pragma unroll kU2
    for (int a = 0; a < 8; a++) {
        if constexpr (kLds) {
            load_op(a, opc_);
        } else if constexpr (kKeep) {
#pragma unroll
            for (int k = 0; k < kNk; k++) opc_[k] = keep[a][k];
        } else {
            if (a < 7) load_op(a + 1, opn_);
        }
        set_phi(a, opc_);
        R f[3] = {0, 0, 0};
        if (kMass) {
            const int nd = nds[a];
#pragma unroll
            for (int k = 0; k < 3; k++) {
                // absolute velocity here: the Rayleigh mass term is not translation invariant
                R al = (R)(C.rdampm * node_v<R>(N, nd, k));
                if (k == 2) al = al + (R)C.gterm;
                f[k] = (R)0 - al * E.emass[blk_at<24>(t, a * 3 + k)];
            }
        }
        const R n1 = opc_[0], n2 = opc_[1], n3 = opc_[2];
        f[0] = f[0] + (n1 * st[0] + n3 * st[4] + n2 * st[5]);
        f[1] = f[1] + (n2 * st[1] + n3 * st[3] + n1 * st[5]);
        f[2] = f[2] + (n3 * st[2] + n2 * st[3] + n1 * st[4]);
        if (kHG == 1) {
            // Kosloff-Frazier stiffness hourglass, src/hrglss.f90:20-55
#pragma unroll
            for (int m = 0; m < 4; m++) {
                const R ph = opc_[3 + m];
#pragma unroll
                for (int k = 0; k < 3; k++) f[k] = f[k] - ph * g[m][k];
            }
        } else if (kHG == 2) {
            // viscous hourglass, src/hrglss.f90:57-97
#pragma unroll
            for (int i = 0; i < 3; i++) {
                R fv = 0;
#pragma unroll
                for (int j = 0; j < 4; j++) fv = fv - coef * qv[i][j] * (R)fi[j][a];
                f[i] = f[i] + fv;
            }
        }
        store_pair_nt(E.ef + reg_ef_at<R>(t, a, 0), f[0], f[1]);
        store_nt(E.ef + reg_ef_at<R>(t, a, 2), f[2]);
        if constexpr (!kLds && !kKeep) {
#pragma unroll
            for (int k = 0; k < kNo; k++) opc_[k] = opn_[k];
        }
    }
}

// ------------------------------------------------------------------------
// PML elements: calcPMLElemKU + calcElemMass + hrglss, one element per thread.
//   src/assembleGlobalKU.f90:36-64, 70-346; src/hrglss.f90
// Streaming like elem_regular_kernel: pass 1 gathers each node once and
// accumulates the velocity gradients and hourglass mode amplitudes; pass 2
// forms each node's 12 (or collapsed 3) force slots, hourglass included, and
// writes every slot once.
// At least 3 waves per SIMD in FP64 (4 in FP32): caps the FP64 per-element
// variant at 168 VGPRs without spills (it would take 176, leaving 2 waves);
// measured 0.324 -> 0.307 ms on the rough-fault box.
template <typename R, int kHG, bool kMass, int kOps>
__global__ void __launch_bounds__(256, (sizeof(R) == 8 ? 3 : (kOps == kOpsGeom ? EQD_PML_GEO_WAVES : 4)))
    elem_pml_kernel(const Member<R> *__restrict__ Ms, const uint16_t *__restrict__ bmem, int remap)
{
    constexpr bool kGeo = kOps == kOpsGeom;
    constexpr bool kLds = kOps == kOpsLds, kCmp = kOps == kOpsCompact || kGeo;
    constexpr int kW = kGeo ? kOpG : (kCmp ? kOpBc : kOpB);
    extern __shared__ __attribute__((aligned(16))) unsigned char eqd_smem[];
    typedef __attribute__((address_space(3))) R lds_R;
    const unsigned gb = launch_block(remap);
    const Member<R> &M = Ms[member_of(bmem, gb)];
    const PmlElems<R> E = M.P;
    const NodeState N = M.N;
    const StepConsts C = M.C;
    const int64_t p = (int64_t)(gb - (unsigned)M.blk0[kKPml]) * blockDim.x + threadIdx.x;
    const R *op = nullptr, *pht = nullptr;
    const lds_R *lrow = nullptr, *lphr = nullptr;
    if constexpr (kLds) {
        // class tables staged in LDS, as in elem_regular_kernel
        lds_R *lop = (lds_R *)eqd_smem, *lph = lop + E.n_opc * kOpW;
        for (int j = threadIdx.x; j < E.n_opc * kOpW; j += blockDim.x) lop[j] = E.optab[j];
        for (int j = threadIdx.x; j < E.n_phic * 32; j += blockDim.x) lph[j] = E.phitab[j];
        __syncthreads();
        if (p >= E.n) return;
        lrow = lop + (int64_t)E.opc[p] * kOpW;
        lphr = lph + (int64_t)E.phic[p] * 32;
    } else {
        if (p >= E.n) return;
        op = E.optab ? E.optab + (int64_t)E.opc[p] * kOpW : nullptr;
        pht = E.phitab ? E.phitab + (int64_t)E.phic[p] * 32 : nullptr;
    }
    const int64_t n = E.n, nn = N.nn;
    constexpr bool kF64 = sizeof(R) == sizeof(double);
    GeoOps<R> geo;
    if constexpr (kGeo)
        geo_ops<R>(E.opb, p, opb_at<R, kOpG>(E.opb, p, kGeoMat), opb_at<R, kOpG>(E.opb, p, kGeoMat + 1), geo);
    auto SHP = [&](int i) -> R {
        if constexpr (kGeo) return geo.shp[i];
        else if constexpr (kLds) return lrow[kOpShp + i];
        else if constexpr (kCmp) return opb_at<R, kW>(E.opb, p, kOpShp + i);
        else return op ? op[kOpShp + i] : opb_at(E.opb, p, kOpShp + i);
    };
    auto PHI = [&](int m, int a) -> R {
        if constexpr (kLds) return lphr[m * 8 + a];
        else return pht ? pht[m * 8 + a] : opb_at(E.opb, p, kOpW + m * 8 + a);
    };
    auto OPK = [&](int k) -> R {
        if constexpr (kGeo) {
            if (k == kOpDet) return geo.det;
            if (k >= kOpSs) return geo.ss[k - kOpSs];
            return opb_at<R, kOpG>(E.opb, p, kGeoMat + (k - kOpLam));
        } else if constexpr (kLds) {
            return lrow[k];
        } else if constexpr (kCmp) {
            return opb_at<R, kW>(E.opb, p, k);
        } else {
            return op ? op[k] : opb_at(E.opb, p, k);
        }
    };
    constexpr bool kCHg = kCmp && kHG == 1;
    R cph[kCHg ? 16 : 1];
    if constexpr (kCHg) {
#pragma unroll
        for (int i = 0; i < 16; i++) {
            if constexpr (kGeo) cph[i] = geo.cph[i];
            else cph[i] = opb_at<R, kW>(E.opb, p, kOpW + i);
        }
    }
    // node ids first, then each node's v and u one node ahead of the arithmetic
    int nds[8];
#pragma unroll
    for (int a = 0; a < 8; a++) nds[a] = load_nt(E.ien + blk_at<8>(p, a));
    constexpr int P = EQD_PREFETCH;
    R vq[P][3], uq[P][3];
#pragma unroll
    for (int b = 0; b < P; b++)
#pragma unroll
        for (int k = 0; k < 3; k++) {
            vq[b][k] = gather_v<R>(N, nds[b], k);
            uq[b][k] = gather_u<R>(N, nds[b], k);
        }
    // velocities (relative to node 1 in the FP32 variant, see elem_regular_kernel)
    R v0[3] = {0, 0, 0}, u0[3] = {0, 0, 0};
    if (!kF64) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            v0[k] = vq[0][k];
            u0[k] = uq[0][k];
        }
    }
    constexpr int fi[4][8] = {{1, 1, -1, -1, -1, -1, 1, 1}, {1, -1, -1, 1, -1, 1, 1, -1},
                              {1, -1, 1, -1, 1, -1, 1, -1}, {1, -1, 1, -1, -1, 1, -1, 1}};
    R sr[6] = {0, 0, 0, 0, 0, 0};
    R Dx_vx = 0, Dy_vy = 0, Dz_vz = 0, Dx_vy = 0, Dy_vx = 0, Dx_vz = 0, Dz_vx = 0, Dy_vz = 0, Dz_vy = 0;
    R pd[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    R qv[3][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}};
    // a node's operator values, one node ahead when they come from memory
    constexpr int kNo = kHG == 1 ? 7 : 3;
    auto set_phi = [&](int a, R o[kNo]) {
        if constexpr (kCHg) {
#pragma unroll
            for (int m = 0; m < 4; m++) o[3 + m] = phi_compact<R>(cph, m, a, o[0], o[1], o[2]);
        }
    };
    auto load_op = [&](int a, R o[kNo]) {
#pragma unroll
        for (int k = 0; k < 3; k++) o[k] = SHP(a * 3 + k);
        if constexpr (kHG == 1 && !kCmp) {
#pragma unroll
            for (int m = 0; m < 4; m++) o[3 + m] = PHI(m, a);
        }
    };
    constexpr bool kKeep = !kLds && !kGeo && (EQD_KEEP_OPS_PML == 2 || (EQD_KEEP_OPS_PML == 1 && sizeof(R) == sizeof(float)));
    constexpr bool kStage = !kLds && !kKeep && !kGeo && EQD_STAGE_SHP_PML;
    constexpr int kNk = kCmp ? 3 : kNo;
    lds_R *stage = (lds_R *)eqd_smem;
    const unsigned bd = blockDim.x, tx = threadIdx.x;
    // the force pass's operator values: dN/dx from the stage, Phi from memory
    // (or rebuilt from the compact operators)
    auto load_op2 = [&](int a, R o[kNo]) {
        if constexpr (kStage) {
#pragma unroll
            for (int k = 0; k < 3; k++) o[k] = stage[(a * 3 + k) * bd + tx];
            if constexpr (kHG == 1 && !kCmp) {
#pragma unroll
                for (int m = 0; m < 4; m++) o[3 + m] = PHI(m, a);
            }
        } else {
            load_op(a, o);
        }
    };
    R opc_[kNo], opn_[kNo];
    R keep[kKeep ? 8 : 1][kNk];
    if constexpr (!kLds) load_op(0, opc_);
constexpr int kU1 = (kKeep || kGeo) ? 8 : EQD_PREFETCH;
#pragma unroll kU1
    for (int a = 0; a < 8; a++) {
        const int b = a % P;
        if constexpr (kLds) {
            load_op(a, opc_);
        } else {
            if constexpr (kKeep) {
#pragma unroll
                for (int k = 0; k < kNk; k++) keep[a][k] = opc_[k];
            }
            set_phi(a, opc_);
            if constexpr (kStage) {
#pragma unroll
                for (int k = 0; k < 3; k++) stage[(a * 3 + k) * bd + tx] = opc_[k];
            }
            if (a < 7) load_op(a + 1, opn_);
        }
        R va[3];
        R dlh[3];
#pragma unroll
        for (int k = 0; k < 3; k++) va[k] = vq[b][k] - v0[k];
        if (kHG == 1) {
#pragma unroll
            for (int k = 0; k < 3; k++) {
                if (kF64)
                    dlh[k] = (R)(uq[b][k] + C.rdampk * (double)va[k]);
                else
                    dlh[k] = (R)((double)(uq[b][k] - u0[k]) + C.rdampk * (double)va[k]);
            }
        }
        if (a + P < 8) {
            const int nx = nds[a + P];
#pragma unroll
            for (int k = 0; k < 3; k++) {
                vq[b][k] = gather_v<R>(N, nx, k);
                if (kHG == 1) uq[b][k] = gather_u<R>(N, nx, k);
            }
        }
        const R n1 = opc_[0], n2 = opc_[1], n3 = opc_[2];
        sr[0] = sr[0] + n1 * va[0];
        sr[1] = sr[1] + n2 * va[1];
        sr[2] = sr[2] + n3 * va[2];
        sr[3] = sr[3] + n3 * va[1] + n2 * va[2];
        sr[4] = sr[4] + n3 * va[0] + n1 * va[2];
        sr[5] = sr[5] + n2 * va[0] + n1 * va[1];
        Dx_vx = Dx_vx + n1 * va[0];
        Dy_vy = Dy_vy + n2 * va[1];
        Dz_vz = Dz_vz + n3 * va[2];
        Dx_vy = Dx_vy + n1 * va[1];
        Dy_vx = Dy_vx + n2 * va[0];
        Dx_vz = Dx_vz + n1 * va[2];
        Dz_vx = Dz_vx + n3 * va[0];
        Dy_vz = Dy_vz + n2 * va[2];
        Dz_vy = Dz_vy + n3 * va[1];
        if (kHG == 1) {
#pragma unroll
            for (int m = 0; m < 4; m++) {
                const R ph = opc_[3 + m];
#pragma unroll
                for (int k = 0; k < 3; k++) pd[m][k] = pd[m][k] + ph * dlh[k];
            }
        } else if (kHG == 2) {
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) qv[i][j] = qv[i][j] + va[i] * (R)fi[j][a];
        }
        if constexpr (!kLds) {
#pragma unroll
            for (int k = 0; k < kNo; k++) opc_[k] = opn_[k];
        }
    }
    const R lam = OPK(kOpLam), miu = OPK(kOpMu), det = OPK(kOpDet);
    const R cdiag = lam + (R)2.0 * miu;
    R srt[6];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        R acc = 0;
#pragma unroll
        for (int j = 0; j < 3; j++) acc = acc + (i == j ? cdiag : lam) * sr[j];
        srt[i] = acc;
    }
#pragma unroll
    for (int i = 3; i < 6; i++) srt[i] = miu * sr[i];

    const R dt = (R)C.dt;
    const R d0 = E.damp[blk_at<3>(p, 0)], d1 = E.damp[blk_at<3>(p, 1)], d2 = E.damp[blk_at<3>(p, 2)];
    const R l2m = lam + (R)2 * miu;
    const R coef[15] = {l2m, lam, lam, lam, l2m, lam, lam, lam, l2m, miu, miu, miu, miu, miu, miu};
    const R grad[15] = {Dx_vx, Dy_vy, Dz_vz, Dx_vx, Dy_vy, Dz_vz, Dx_vx, Dy_vy, Dz_vz,
                        Dx_vy, Dy_vx, Dx_vz, Dz_vx, Dy_vz, Dz_vy};
    constexpr int dir[15] = {0, 1, 2, 0, 1, 2, 0, 1, 2, 0, 1, 0, 2, 1, 2};
    R sv[15];
#pragma unroll
    for (int c = 0; c < 15; c++) {
        const R dd = dir[c] == 0 ? d0 : (dir[c] == 1 ? d1 : d2);
        R sc = load_nt(E.stress + blk_at<21>(p, c));
        sc = coef[c] * grad[c] + ((R)1 / dt - dd / (R)2) * sc;
        sc = sc / ((R)1 / dt + dd / (R)2);
        sv[c] = sc;
        store_nt(E.stress + blk_at<21>(p, c), sc);
    }
    const R sxx = sv[0] + sv[1] + sv[2];
    const R syy = sv[3] + sv[4] + sv[5];
    const R szz = sv[6] + sv[7] + sv[8];
    const R sxy = sv[9] + sv[10];
    const R sxz = sv[11] + sv[12];
    const R syz = sv[13] + sv[14];
    // the summed stresses, for the PML nodes' split forces (node_pml_kernel)
    store_pair(E.ssum + pml_at<R, 6>(p, 0), sxx, syy);
    store_pair(E.ssum + pml_at<R, 6>(p, 2), szz, sxy);
    store_pair(E.ssum + pml_at<R, 6>(p, 4), sxz, syz);
    R s0[6];
#pragma unroll
    for (int k = 0; k < 6; k++) s0[k] = load_nt(E.stress + blk_at<21>(p, 15 + k)) + (R)C.rdampk * srt[k];
    const R dw = det * (R)C.w;
    const unsigned mask = E.mask12[p];
    R g[4][3];
    R coef2 = 0;
    if (kHG == 1) {
        R ss[6];
#pragma unroll
        for (int i = 0; i < 6; i++) ss[i] = OPK(kOpSs + i);
#pragma unroll
        for (int m = 0; m < 4; m++) {
            g[m][0] = ss[0] * pd[m][0] + ss[1] * pd[m][1] + ss[2] * pd[m][2];
            g[m][1] = ss[1] * pd[m][0] + ss[3] * pd[m][1] + ss[4] * pd[m][2];
            g[m][2] = ss[2] * pd[m][0] + ss[4] * pd[m][1] + ss[5] * pd[m][2];
        }
    } else if (kHG == 2) {
        const R rho = OPK(kOpRho), vp = OPK(kOpVp);
        coef2 = (R)0.25 * (R)C.kapa_hg * rho * vp * (R)pow((double)det * C.w, 2.0 / 3.0);
    }

    if constexpr (!kLds && !kKeep) load_op2(0, opc_);
constexpr int kU2 = (kKeep || kGeo) ? 8 : EQD_ELEM_UNROLL;
#pragma unroll kU2
    for (int a = 0; a < 8; a++) {
        if constexpr (kLds) {
            load_op(a, opc_);
        } else if constexpr (kKeep) {
#pragma unroll
            for (int k = 0; k < kNk; k++) opc_[k] = keep[a][k];
        } else {
            if (a < 7) load_op2(a + 1, opn_);
        }
        set_phi(a, opc_);
        // hourglass force on this node, subtracted mode by mode
        R hg[3] = {0, 0, 0};
        if (kHG == 1) {
#pragma unroll
            for (int m = 0; m < 4; m++) {
                const R ph = opc_[3 + m];
#pragma unroll
                for (int k = 0; k < 3; k++) hg[k] = hg[k] - ph * g[m][k];
            }
        } else if (kHG == 2) {
#pragma unroll
            for (int i = 0; i < 3; i++) {
                R fv = 0;
#pragma unroll
                for (int j = 0; j < 4; j++) fv = fv - coef2 * qv[i][j] * (R)fi[j][a];
                hg[i] = fv;
            }
        }
        const R g1 = opc_[0], g2 = opc_[1], g3 = opc_[2];
        R f[12];
        if (kMass) {
            const int nd = nds[a];
#pragma unroll
            for (int k = 0; k < 9; k++) f[k] = 0;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                R al = (R)(C.rdampm * node_v<R>(N, nd, k));
                if (k == 2) al = al + (R)C.gterm;
                f[9 + k] = (R)0 - al * E.emass[blk_at<24>(p, a * 3 + k)];
            }
        } else {
#pragma unroll
            for (int k = 0; k < 12; k++) f[k] = 0;
        }
        f[0] = f[0] - dw * g1 * sxx;
        f[1] = f[1] - dw * g2 * sxy;
        f[2] = f[2] - dw * g3 * sxz;
        f[3] = f[3] - dw * g1 * sxy;
        f[4] = f[4] - dw * g2 * syy;
        f[5] = f[5] - dw * g3 * syz;
        f[6] = f[6] - dw * g1 * sxz;
        f[7] = f[7] - dw * g2 * syz;
        f[8] = f[8] - dw * g3 * szz;
        f[9] = f[9] - dw * (g1 * s0[0] + g3 * s0[4] + g2 * s0[5]);
        f[10] = f[10] - dw * (g2 * s0[1] + g3 * s0[3] + g1 * s0[5]);
        f[11] = f[11] - dw * (g3 * s0[2] + g2 * s0[3] + g1 * s0[4]);
        if (mask & (1u << a)) {
            // a PML node: its 9 split components are (0 - (dw g_i) s_j) with
            // the summed stresses stored above, so the node kernel forms them
            // from dw g (3 values) bit for bit; 10-12 are stored whole
            store_pair_nt(E.ef + pml_at<R, kEfPml>(p, a * 6 + 0), dw * g1, dw * g2);
            store_pair_nt(E.ef + pml_at<R, kEfPml>(p, a * 6 + 2), dw * g3, f[9] + hg[0]);
            store_pair_nt(E.ef + pml_at<R, kEfPml>(p, a * 6 + 4), f[10] + hg[1], f[11] + hg[2]);
        } else {
            // a 3-component node takes the collapsed sums (src/assembleGlobalKU.f90:55-61)
            store_pair_nt(E.ef + pml_at<R, kEfPml>(p, a * 6 + 0), f[0] + f[1] + f[2] + f[9] + hg[0],
                          f[3] + f[4] + f[5] + f[10] + hg[1]);
            store_nt(E.ef + pml_at<R, kEfPml>(p, a * 6 + 2), f[6] + f[7] + f[8] + f[11] + hg[2]);
        }
        if constexpr (!kLds && !kKeep) {
#pragma unroll
            for (int k = 0; k < kNo; k++) opc_[k] = opn_[k];
        }
    }
}

// ------------------------------------------------------------------------
// Gather of one 3-component node: sum of its elements' forces in ascending
// element order (assembleGlobalKU's scatter-add, src/assembleGlobalKU.f90:28-63).
template <typename R>
__device__ __forceinline__ void gather3(const NodeState &N, int node, const R *ef, const R *efp, double F[3])
{
    // entries in batches of eight: all entry loads, then all force loads,
    // then the sums in entry order (so at most two memory latencies per batch)
    F[0] = 0.0; F[1] = 0.0; F[2] = 0.0;
    const int beg = N.rowptr[node], end = N.rowptr[node + 1];
    for (int j0 = beg; j0 < end; j0 += 8) {
        uint32_t e[8];
#pragma unroll
        for (int k = 0; k < 8; k++) e[k] = j0 + k < end ? N.ent[j0 + k] : 0u;
        R f[8][3];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            if (j0 + k < end) {
                const int64_t idx = e[k] & kEntIdxMask;
                const int a = (e[k] >> kEntLocalShift) & 7;
                if (e[k] & kEntPml) {
                    load_pair(efp + pml_at<R, kEfPml>(idx, a * 6), f[k][0], f[k][1]);
                    f[k][2] = efp[pml_at<R, kEfPml>(idx, a * 6 + 2)];
                } else {
                    load_pair(ef + reg_ef_at<R>(idx, a, 0), f[k][0], f[k][1]);
                    f[k][2] = ef[reg_ef_at<R>(idx, a, 2)];
                }
            }
        }
#pragma unroll
        for (int k = 0; k < 8; k++) {
            if (j0 + k < end) {
                F[0] = F[0] + (double)f[k][0];
                F[1] = F[1] + (double)f[k][1];
                F[2] = F[2] + (double)f[k][2];
            }
        }
    }
}

// gather3 from the slot-major table (entry k of free node i at
// blk_at<kEll3>(i, k), unused slots kEntNone after the used ones): the same
// loads and the same sums in the same order as gather3
template <typename R>
__device__ __forceinline__ void gather3_ell(const uint32_t *ell, int64_t i, const R *ef, const R *efp, double F[3])
{
    F[0] = 0.0; F[1] = 0.0; F[2] = 0.0;
    uint32_t e[kEll3];
#pragma unroll
    for (int k = 0; k < kEll3; k++) e[k] = ell[blk_at<kEll3>(i, k)];
    R f[kEll3][3];
#pragma unroll
    for (int k = 0; k < kEll3; k++) {
        if (e[k] != kEntNone) {
            const int64_t idx = e[k] & kEntIdxMask;
            const int a = (e[k] >> kEntLocalShift) & 7;
            if (e[k] & kEntPml) {
                load_pair(efp + pml_at<R, kEfPml>(idx, a * 6), f[k][0], f[k][1]);
                f[k][2] = efp[pml_at<R, kEfPml>(idx, a * 6 + 2)];
            } else {
                load_pair(ef + reg_ef_at<R>(idx, a, 0), f[k][0], f[k][1]);
                f[k][2] = ef[reg_ef_at<R>(idx, a, 2)];
            }
        }
    }
#pragma unroll
    for (int k = 0; k < kEll3; k++) {
        if (e[k] != kEntNone) {
            F[0] = F[0] + (double)f[k][0];
            F[1] = F[1] + (double)f[k][1];
            F[2] = F[2] + (double)f[k][2];
        }
    }
}

// First NaN velocity: the reference stops at the first node, in node order, of
// the first step whose update gives a NaN (src/driver.f90:147-152).  The key
// (step << 32 | node + 1) is lowered atomically, so the earliest step and then
// the lowest node wins whatever thread or stream finds it first.
__device__ __forceinline__ void flag_nan(const NodeState &N, int node, int step, double v0, double v1, double v2)
{
    if (v0 != v0 || v1 != v1 || v2 != v2) {
        const unsigned long long key = ((unsigned long long)(uint32_t)step << 32) | (uint32_t)(node + 1);
        atomicMin(N.nanflag, key);
    }
}

// Node kernel for 3-component nodes that are not split nodes.
// mode 0: gather, a = F/m, then the update of the next step (fused);
// mode 1: gather, store a = F/m (the last step of a run).
template <typename R>
__global__ void __launch_bounds__(256) node3_kernel(const Member<R> *__restrict__ Ms, const uint16_t *__restrict__ bmem,
                                                    int remap)
{
    const unsigned gb = launch_block(remap);
    const Member<R> &M = Ms[member_of(bmem, gb)];
    const int64_t i = (int64_t)(gb - (unsigned)M.blk0[kKNode3]) * blockDim.x + threadIdx.x;
    if (i >= M.n_node3_free) return;
    const NodeState N = M.N;
    const R *ef = M.E.ef, *efp = M.P.ef;
    const double dt = M.C.dt;
    const Clock ck = *M.clk;
    const int mode = ck.mode, next_step = ck.nt + 1;
    const int node = M.node3_free[i];
    const int64_t nn = N.nn;
    double F[3];
    if (M.ell3) gather3_ell<R>(M.ell3, i, ef, efp, F);
    else gather3<R>(N, node, ef, efp, F);
    const double m = node_mass<R>(N, node);
    if (mode == 1) {
#pragma unroll
        for (int k = 0; k < 3; k++) N.acc[blk_at<3>(node, k)] = F[k] / m;
        return;
    }
    double vn[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const double a = F[k] / m;
        const double v = node_v<R>(N, node, k) + a * dt;
        node_store<R>(N, node, k, v, node_u<R>(N, node, k) + v * dt);
        vn[k] = v;
    }
    flag_nan(N, node, next_step, vn[0], vn[1], vn[2]);
}

// 12-component (PML) nodes; the gather reads only PML element forces because
// regular elements never touch a PML node (checked on the host).
// src/driver.f90:107-141 for the damped split update.
template <typename R>
__global__ void __launch_bounds__(256) node_pml_kernel(const Member<R> *__restrict__ Ms,
                                                       const uint16_t *__restrict__ bmem, int remap)
{
    const unsigned gb = launch_block(remap);
    const Member<R> &M = Ms[member_of(bmem, gb)];
    const int64_t q = (int64_t)(gb - (unsigned)M.blk0[kKNodePml]) * blockDim.x + threadIdx.x;
    if (q >= M.PN.n) return;
    const NodeState N = M.N;
    const PmlNodes P = M.PN;
    const R *efp = M.P.ef, *ssum = M.P.ssum;
    const double dt = M.C.dt;
    const Clock ck = *M.clk;
    const int mode = ck.mode, next_step = ck.nt + 1;
    const int node = P.node[q];
    const int64_t nn = N.nn;
    double F[12];
#pragma unroll
    for (int k = 0; k < 12; k++) F[k] = 0.0;
    // entries in batches of four: entry loads, force loads, then the sums in order
    const int beg = N.rowptr[node], end = N.rowptr[node + 1];
    for (int j0 = beg; j0 < end; j0 += 4) {
        uint32_t e[4];
#pragma unroll
        for (int k = 0; k < 4; k++) e[k] = j0 + k < end ? N.ent[j0 + k] : 0u;
        R f[4][12];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (j0 + k < end) {
                const int64_t idx = e[k] & kEntIdxMask;
                const int a = (e[k] >> kEntLocalShift) & 7;
                // the 9 split components from dw g and the element's summed
                // stresses, as elem_pml_kernel forms them; 10-12 as stored
                R q1, q2, q3, d1, d2, d3, sxx, syy, szz, sxy, sxz, syz;
                load_pair(efp + pml_at<R, kEfPml>(idx, a * 6 + 0), q1, q2);
                load_pair(efp + pml_at<R, kEfPml>(idx, a * 6 + 2), q3, d1);
                load_pair(efp + pml_at<R, kEfPml>(idx, a * 6 + 4), d2, d3);
                load_pair(ssum + pml_at<R, 6>(idx, 0), sxx, syy);
                load_pair(ssum + pml_at<R, 6>(idx, 2), szz, sxy);
                load_pair(ssum + pml_at<R, 6>(idx, 4), sxz, syz);
                f[k][0] = (R)0 - q1 * sxx;
                f[k][1] = (R)0 - q2 * sxy;
                f[k][2] = (R)0 - q3 * sxz;
                f[k][3] = (R)0 - q1 * sxy;
                f[k][4] = (R)0 - q2 * syy;
                f[k][5] = (R)0 - q3 * syz;
                f[k][6] = (R)0 - q1 * sxz;
                f[k][7] = (R)0 - q2 * syz;
                f[k][8] = (R)0 - q3 * szz;
                f[k][9] = d1;
                f[k][10] = d2;
                f[k][11] = d3;
            }
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (j0 + k < end) {
#pragma unroll
                for (int c = 0; c < 12; c++) F[c] = F[c] + (double)f[k][c];
            }
        }
    }
    const double m = node_mass<R>(N, node);
    if (mode == 1) {
#pragma unroll
        for (int k = 0; k < 12; k++) P.acc[blk_at<12>(q, k)] = F[k] / m;
        return;
    }
    const double d[3] = {pml_damp<R>(P, q, 0), pml_damp<R>(P, q, 1), pml_damp<R>(P, q, 2)};
    double v[12];
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const double dv = d[k % 3];
        v[k] = (F[k] / m + pml_v<R>(P, q, k) * (1.0 / dt - dv / 2.0)) / (1.0 / dt + dv / 2.0);
    }
#pragma unroll
    for (int k = 9; k < 12; k++) v[k] = pml_v<R>(P, q, k) + F[k] / m * dt;
#pragma unroll
    for (int k = 0; k < 12; k++) pml_store<R>(P, q, k, v[k]);
    double vel[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        vel[k] = v[3 * k] + v[3 * k + 1] + v[3 * k + 2] + v[9 + k];
        node_store<R>(N, node, k, vel[k], node_u<R>(N, node, k) + vel[k] * dt);
    }
    flag_nan(N, node, next_step, vel[0], vel[1], vel[2]);
}

// Pending update at the start of a run: v += a*dt from the stored F/m
// (velDispUpdate with the previous step's nodalForceArr).
template <typename R>
__global__ void __launch_bounds__(256) update3_from_acc_kernel(const Member<R> *__restrict__ Ms,
                                                               const uint16_t *__restrict__ bmem)
{
    const unsigned gb = blockIdx.x;
    const Member<R> &M = Ms[member_of(bmem, gb)];
    const int64_t i = (int64_t)(gb - (unsigned)M.blk0[kKPend3]) * blockDim.x + threadIdx.x;
    if (i >= M.n_node3_all) return;
    const NodeState N = M.N;
    const double dt = M.C.dt;
    const int next_step = M.clk->nt + 1;
    const int node = M.node3_all[i];
    const int64_t nn = N.nn;
    double vn[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const double v = exact_v(M, node, k) + N.acc[blk_at<3>(node, k)] * dt;
        exact_store(M, node, k, v, exact_u(M, node, k) + v * dt);
        vn[k] = v;
    }
    flag_nan(N, node, next_step, vn[0], vn[1], vn[2]);
}

template <typename R>
__global__ void __launch_bounds__(256) update_pml_from_acc_kernel(const Member<R> *__restrict__ Ms,
                                                                  const uint16_t *__restrict__ bmem)
{
    const unsigned gb = blockIdx.x;
    const Member<R> &M = Ms[member_of(bmem, gb)];
    const int64_t q = (int64_t)(gb - (unsigned)M.blk0[kKPendPml]) * blockDim.x + threadIdx.x;
    if (q >= M.PN.n) return;
    const NodeState N = M.N;
    const PmlNodes P = M.PN;
    const double dt = M.C.dt;
    const int next_step = M.clk->nt + 1;
    const int node = P.node[q];
    const int64_t nn = N.nn;
    const double d[3] = {pml_damp<R>(P, q, 0), pml_damp<R>(P, q, 1), pml_damp<R>(P, q, 2)};
    double v[12];
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const double dv = d[k % 3];
        v[k] = (P.acc[blk_at<12>(q, k)] + pml_v<R>(P, q, k) * (1.0 / dt - dv / 2.0)) / (1.0 / dt + dv / 2.0);
    }
#pragma unroll
    for (int k = 9; k < 12; k++) v[k] = pml_v<R>(P, q, k) + P.acc[blk_at<12>(q, k)] * dt;
#pragma unroll
    for (int k = 0; k < 12; k++) pml_store<R>(P, q, k, v[k]);
    double vel[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        vel[k] = v[3 * k] + v[3 * k + 1] + v[3 * k + 2] + v[9 + k];
        node_store<R>(N, node, k, vel[k], node_u<R>(N, node, k) + vel[k] * dt);
    }
    flag_nan(N, node, next_step, vel[0], vel[1], vel[2]);
}

// ------------------------------------------------------------------------
// Friction laws, src/fric.f90
__device__ __forceinline__ double slip_weak(double slip, const double *f, int64_t nf, int64_t i)
{
    const double fs = f[0 * nf + i], fd = f[1 * nf + i], d0 = f[2 * nf + i];
    double xmu = 0.0;
    if (fabs(slip) < 1.0e-10)
        xmu = fs;
    else if (slip < d0)
        xmu = fs - (fs - fd) * slip / d0;
    if (slip >= d0) xmu = fd;
    return xmu;
}

__device__ __forceinline__ double time_weak(double trupt, const double *f, int64_t nf, int64_t i)
{
    const double fs = f[0 * nf + i], fd = f[1 * nf + i], t0 = f[4 * nf + i];
    if (trupt <= 0.0) return fs;
    if (trupt < t0) return fs - (fs - fd) * trupt / t0;
    return fd;
}

struct RsfParams {
    double A, B, L, f0, V0, fw, Vw;
};

// rate_state_ageing_law, src/fric.f90:39-61
__device__ __forceinline__ void ageing_law(double V2, double &theta, const RsfParams &P, double dt, double &xmu,
                                           double &dxmudv)
{
    const double tmpc = 1.0 / (2.0 * P.V0) * exp((P.f0 + P.B * log(P.V0 * theta / P.L)) / P.A);
    const double tmp = (V2 + 1.0e-30) * tmpc;
    xmu = P.A * log(tmp + sqrt(tmp * tmp + 1.0));
    dxmudv = P.A * tmpc / sqrt(1.0 + tmp * tmp);
    theta = P.L / V2 + (theta - P.L / V2) * exp(-V2 * dt / P.L);
}

// rate_state_slip_law, src/fric.f90:63-95
__device__ __forceinline__ void slip_law(double V2, double &psi, const RsfParams &P, double dt, double &xmu,
                                         double &dxmudv)
{
    const double tmpc = 1.0 / (2.0 * P.V0) * exp(psi / P.A);
    const double tmp = (V2 + 1.0e-30) * tmpc;
    xmu = P.A * log(tmp + sqrt(tmp * tmp + 1.0));
    dxmudv = P.A * tmpc / sqrt(1.0 + tmp * tmp);
    const double fLV = P.f0 - (P.B - P.A) * log(V2 / P.V0);
    const double r = V2 / P.Vw;
    double r8 = r * r;
#pragma unroll
    for (int k = 2; k < 8; k++) r8 = r * r8;
    const double fss = P.fw + (fLV - P.fw) / pow(1.0 + r8, 0.125);
    const double fssa = fss / P.A;
    const double psiss = P.A * log(2.0 * P.V0 / V2 * (exp(fssa) - exp(-fssa)) / 2.0);
    psi = psiss + (psi - psiss) * exp(-V2 * dt / P.L);
}

// ------------------------------------------------------------------------
// Split-node fault, one node pair per thread: getNsdSlipSliprateTraction,
// solveSWTW / solveRSF + NewtonRaphson, on-fault stations, rupture time
// (src/faulting.f90:3-541), then F/m and (mode 0) the next step's update of
// the pair's two nodes.
//
// kLearn: the rate-and-state Newton starts from the learned guess (one
// wavefront evaluates the network for its 64 pairs, so every lane stays until
// then: lanes past the last pair read pair nf-1 and write only their own
// padding rows of fric, then leave).
template <typename R, bool kLearn>
__global__ void __launch_bounds__(128, EQD_FAULT_WAVES) fault_kernel(const Member<R> *__restrict__ Ms,
                                                                      const uint16_t *__restrict__ bmem)
{
    __shared__ __attribute__((aligned(16))) __bf16 mlp_lds0[kLearn ? 2 * 64 * kLd0 : 8];
    __shared__ __attribute__((aligned(16))) __bf16 mlp_lds[kLearn ? 2 * 64 * kLd : 8];
    const unsigned gb = blockIdx.x;
    const Member<R> &M = Ms[member_of(bmem, gb)];
    const FaultState Fs = M.F;
    const NodeState N = M.N;
    const StepConsts C = M.C;
    const R *ef = M.E.ef, *efp = M.P.ef;
    const Clock ck = *M.clk;
    const int nt = ck.nt, mode = ck.mode;
    const double time = ck.time;
    const int64_t i0 = (int64_t)(gb - (unsigned)M.blk0[kKFault]) * blockDim.x + threadIdx.x;
    if (!kLearn && i0 >= Fs.nf) return;
    const bool active = i0 < Fs.nf;
    const int64_t i = active ? i0 : Fs.nf - 1;
    const int64_t nf = Fs.nf, nn = N.nn, ld = Fs.ld;
    double *fr = Fs.fric;
#define FR(k) fr[(int64_t)((k) - 1) * ld + i0]
    const double dt = C.dt, Ce = (double)C.C_elastic;
    const int ns = Fs.nsmp[i], nm = Fs.nsmp[nf + i];
    // the pair's normals and its nodes' velocity and displacement are read
    // again where they are needed after the friction solve (no other kernel
    // writes split nodes), so they do not stay live across the Newton loop
    auto UN = [&](int j) { return Fs.un[j * nf + i]; };
    auto US = [&](int j) { return Fs.us[j * nf + i]; };
    auto UD = [&](int j) { return Fs.ud[j * nf + i]; };
    const double arn = Fs.arn[i];
    // velocity and displacement of the pair's nodes (side 0 slave, 1 master)
    // at full precision: the FP32 bulk mode keeps them in the fault's FP64
    // arrays, the FP64 mode in the node arrays
    auto PV = [&](int side, int node, int k) -> double {
        if constexpr (sizeof(R) == sizeof(float)) return Fs.sv[(side * 3 + k) * ld + i];
        else return N.vel[blk_at<3>(node, k)];
    };
    auto PU = [&](int side, int node, int k) -> double {
        if constexpr (sizeof(R) == sizeof(float)) return Fs.su[(side * 3 + k) * ld + i];
        else return N.disp[blk_at<3>(node, k)];
    };
    double Fsl[3], Fma[3];
    gather3<R>(N, ns, ef, efp, Fsl);
    gather3<R>(N, nm, ef, efp, Fma);
    double nsd[3][2][3];
    {
    const double un[3] = {UN(0), UN(1), UN(2)}, us[3] = {US(0), US(1), US(2)}, ud[3] = {UD(0), UD(1), UD(2)};
    double q[3][2][3];  // [xyz][slave/master][force, vel, disp]
    for (int k = 0; k < 3; k++) {
        q[k][0][0] = Fsl[k];
        q[k][1][0] = Fma[k];
        q[k][0][1] = PV(0, ns, k);
        q[k][1][1] = PV(1, nm, k);
        q[k][0][2] = PU(0, ns, k);
        q[k][1][2] = PU(1, nm, k);
    }
    for (int j = 0; j < 3; j++)
        for (int s = 0; s < 2; s++) {
            nsd[0][s][j] = q[0][s][j] * un[0] + q[1][s][j] * un[1] + q[2][s][j] * un[2];
            nsd[1][s][j] = q[0][s][j] * us[0] + q[1][s][j] * us[1] + q[2][s][j] * us[2];
            nsd[2][s][j] = q[0][s][j] * ud[0] + q[1][s][j] * ud[1] + q[2][s][j] * ud[2];
        }
    }
    double slip[4], rate[4], T[4], T0[3];
    T0[0] = FR(7);
    T0[1] = FR(8) + 0.0;
    T0[2] = FR(49);
    const double ms = Fs.fms[i], mm = Fs.fmm[i];
    const double totalMass = (ms + mm) * arn;
    for (int j = 0; j < 3; j++) slip[j] = nsd[j][1][2] - nsd[j][0][2];
    slip[3] = sqrt(slip[0] * slip[0] + slip[1] * slip[1] + slip[2] * slip[2]);
    for (int j = 0; j < 3; j++) rate[j] = nsd[j][1][1] - nsd[j][0][1];
    rate[3] = sqrt(rate[0] * rate[0] + rate[1] * rate[1] + rate[2] * rate[2]);
    FR(71) = slip[1];
    FR(72) = slip[2];
    FR(73) = slip[0];
    FR(74) = rate[1];
    FR(75) = rate[2];
    if (rate[3] > FR(76)) FR(76) = rate[3];
    FR(77) = FR(77) + rate[3] * dt;
    T[0] = (ms * mm * ((nsd[0][1][1] - nsd[0][0][1]) + (nsd[0][1][2] - nsd[0][0][2]) / dt) / dt + ms * nsd[0][1][0] -
            mm * nsd[0][0][0]) / totalMass + T0[0] * Ce;
    T[1] = (ms * mm * (nsd[1][1][1] - nsd[1][0][1]) / dt + ms * nsd[1][1][0] - mm * nsd[1][0][0]) / totalMass + T0[1] * Ce;
    T[2] = (ms * mm * (nsd[2][1][1] - nsd[2][0][1]) / dt + ms * nsd[2][1][0] - mm * nsd[2][0][0]) / totalMass + T0[2] * Ce;

    const double sx = Fs.xyz[i], sy = Fs.xyz[nf + i], sz = Fs.xyz[2 * nf + i];
    if (C.friclaw >= 3 && C.C_nuclea == 1 && C.nucfault == 1) {
        // rsfNucleation, src/faulting.f90:367-414
        double dtau = 0.0;
        const double radius = sqrt((sx - C.xsource) * (sx - C.xsource) + (sy - C.ysource) * (sy - C.ysource) +
                                   (sz - C.zsource) * (sz - C.zsource));
        double Fn = 0.0, G = 1.0;
        if (radius < C.nucR) Fn = exp(radius * radius / (radius * radius - C.nucR * C.nucR));
        if (time <= C.nucT) G = exp((time - C.nucT) * (time - C.nucT) / (time * (time - 2.0 * C.nucT)));
        if (C.TPV == 105 || C.TPV == 104) {
            dtau = C.nucdtau0 * Fn * G;
        } else if (C.TPV == 2802) {
            if (nt == 1) {
                FR(81) = C.nucdtau0;
                const double ttao = sqrt(T[1] * T[1] + T[2] * T[2]);
                const double back = sqrt((rate[1] + FR(26)) * (rate[1] + FR(26)) + (rate[2] + FR(27)) * (rate[2] + FR(27)));
                FR(20) = FR(9) * log(2.0 * FR(12) / back * sinh(ttao / fabs(T[0]) / FR(9)));
                FR(23) = fabs(T[0]);
            }
            dtau = FR(81) * Fn * G;
        }
        T[1] = T[1] + dtau;
    }
    T[3] = sqrt(T[1] * T[1] + T[2] * T[2]);

    if (C.friclaw <= 2) {
        // solveSWTW, src/faulting.f90:136-188
        double mu = 0.0;
        if (C.friclaw == 1)
            mu = slip_weak(FR(77), fr, ld, i0);
        else
            mu = time_weak(time - Fs.fnft[i], fr, ld, i0);
        if (C.C_nuclea == 1 && C.nucfault == 1) {
            // swtwNucleation, src/faulting.f90:416-443
            const double radius = sqrt((sx - C.xsource) * (sx - C.xsource) + (sy - C.ysource) * (sy - C.ysource) +
                                       (sz - C.zsource) * (sz - C.zsource));
            double tr = 1.0e9, tc;
            if (radius <= C.nucR) {
                if (C.TPV == 201 || C.TPV == 36 || C.TPV == 37)
                    tr = (radius + 0.081 * C.nucR * (1.0 / (1.0 - (radius / C.nucR) * (radius / C.nucR)) - 1.0)) /
                         (0.7 * 3464.0);
                if (C.TPV == 202) tr = radius / C.nucRuptVel;
            }
            tc = 1.0;
            if (time < tr)
                tc = 0.0;
            else if ((time < (tr + FR(5))) && (time >= tr))
                tc = (time - tr) / FR(5);
            const double cand = FR(1) + (FR(2) - FR(1)) * tc;
            mu = cand < mu ? cand : mu;
        }
        const double eff = (T[0] + FR(6)) > 0 ? 0.0 : T[0] + FR(6);
        const double trialShear = FR(4) - mu * eff;
        if (T[3] > trialShear) {
            T[1] = T[1] * trialShear / T[3];
            T[2] = T[2] * trialShear / T[3];
        }
        for (int j = 0; j < 3; j++) {
            const double xT = (T[0] * UN(j) + T[1] * US(j) + T[2] * UD(j)) * arn;
            const double xT0 = (T0[0] * UN(j) + T0[1] * US(j) + T0[2] * UD(j)) * arn;
            Fsl[j] = Fsl[j] + xT - xT0 * Ce;
            Fma[j] = Fma[j] - xT + xT0 * Ce;
        }
        FR(78) = T[0];
        FR(79) = T[1];
        FR(80) = T[2];
    } else {
        // solveRSF + NewtonRaphson, src/faulting.f90:190-327, 459-516
        const RsfParams P = {FR(9), FR(10), FR(11), FR(13), FR(12), FR(14), FR(15)};
        double trial[4];
        if (C.friclaw == 5)
            T[0] = T[0] + FR(51);
        else
            T[0] = T[0] + FR(6);
        if (C.insertFaultType > 0 && C.C_elastic == 1) {
            const double max_norm = -40.0e6, min_norm = -10.0e6;
            if (T[0] >= min_norm)
                T[0] = min_norm;
            else if (T[0] <= max_norm)
                T[0] = max_norm;
        }
        if (T[0] > 0.0) T[0] = 0.0;
        for (int j = 0; j < 3; j++) {
            slip[j] = slip[j] + FR(25 + j) * time;
            rate[j] = rate[j] + FR(25 + j);
        }
        slip[3] = sqrt(slip[1] * slip[1] + slip[2] * slip[2]);
        rate[3] = sqrt(rate[1] * rate[1] + rate[2] * rate[2]);
        double v_trial = rate[3];
        double theta_pc_tmp = FR(23);
        // rate_state_normal_stress, src/faulting.f90:35-52
        double theta_pc_dot = -(v_trial / P.L * (FR(23) - fabs(T[0])));
        FR(23) = FR(23) + theta_pc_dot * dt;
        FR(24) = theta_pc_dot;
        double statetmp = FR(20);
        double xmu, dxmudv, st20 = FR(20);
        if (C.friclaw == 3)
            ageing_law(v_trial, st20, P, dt, xmu, dxmudv);
        else
            slip_law(v_trial, st20, P, dt, xmu, dxmudv);
        FR(20) = st20;
        const double taoc_old = C.friclaw == 5 ? FR(4) - xmu * T[0] : xmu * theta_pc_tmp;
        const double mr = mm * ms / (mm + ms);
        const double Tc = arn * dt / mr;
        for (int j = 1; j <= 2; j++) trial[j] = T[j] - taoc_old * 0.5 * (rate[j] / rate[3]) + FR(25 + j) / Tc;
        trial[3] = sqrt(trial[1] * trial[1] + trial[2] * trial[2]);
        double taoc_new = 0.0, stateTmp = 0.0, thetaPcTmp = 0.0;
        // NewtonRaphson, src/faulting.f90:459-516, from v0 for at most maxit
        // iterations; 1 when the convergence test was met, 0 if not; its =
        // evaluations.  With roundoff_stop, also 2 when the residual is within
        // 16 ulps of the size of its terms (no iterate can then do better).
        auto newton = [&](double v0, int maxit, int &its, bool roundoff_stop) -> int {
            v_trial = v0;
            for (int iv = 1; iv <= maxit; iv++) {
                its = iv;
                double rsfeq, drsfeqdv;
                stateTmp = statetmp;
                if (C.friclaw == 3)
                    ageing_law(v_trial, stateTmp, P, dt, xmu, dxmudv);
                else
                    slip_law(v_trial, stateTmp, P, dt, xmu, dxmudv);
                if (C.friclaw < 5) {
                    thetaPcTmp = theta_pc_tmp;
                    theta_pc_dot = -(v_trial / P.L * (thetaPcTmp - fabs(T[0])));
                    thetaPcTmp = thetaPcTmp + theta_pc_dot * dt;
                    taoc_new = xmu * thetaPcTmp;
                    rsfeq = v_trial + Tc * (taoc_new * 0.5 - trial[3]);
                    drsfeqdv = 1.0 + Tc * (dxmudv * thetaPcTmp) * 0.5;
                } else {
                    const double tn = T[0] < 0.0 ? T[0] : 0.0;
                    taoc_new = FR(4) - xmu * tn;
                    rsfeq = v_trial + Tc * (taoc_new * 0.5 - trial[3]);
                    drsfeqdv = 1.0 + Tc * (-dxmudv * tn) * 0.5;
                }
                if (fabs(rsfeq / drsfeqdv) < 1.0e-14 * fabs(v_trial) && fabs(rsfeq) < 1.0e-6 * fabs(v_trial))
                    return 1;
                if (roundoff_stop &&
                    fabs(rsfeq) <= 16.0 * 2.220446049250313e-16 *
                                       (fabs(v_trial) + Tc * (0.5 * fabs(taoc_new) + fabs(trial[3]))))
                    return 2;
                const double nv = v_trial - rsfeq / drsfeqdv;
                if (nv <= 0.0)
                    v_trial = v_trial / 2.0;
                else
                    v_trial = nv;
            }
            return 0;
        };
        // inputs of the learned guess (and of a harvested training row)
        const double v_plain = v_trial;
        auto features = [&](float feat[kMlpIn]) {
            feat[0] = (float)trial[3];
            feat[1] = (float)T[0];
            feat[2] = (float)(C.friclaw == 3 ? log(statetmp) : statetmp);
            feat[3] = (float)P.A;
            feat[4] = (float)P.B;
            feat[5] = (float)log(v_plain > 1.0e-30 ? v_plain : 1.0e-30);
        };
        int its = 0, its2 = 0, fb = 0, res;
        if (Fs.k_fallback > 0) {
            // iteration policy of the faster mode: start from the proposal (the
            // network's where it predicts a change, else the plain guess), stop
            // at convergence or at the roundoff floor, and fall back to the
            // reference's solve from the plain guess after k_fallback iterations
            double v0 = v_plain;
            if constexpr (kLearn) {
                const int w = (int)(threadIdx.x >> 6);
                float feat[kMlpIn];
                features(feat);
                const float y = mlp_wave_forward(Fs.net, feat, mlp_lds0 + w * 64 * kLd0, mlp_lds + w * 64 * kLd);
                if (!active) return;
                const double r = (double)(y * Fs.net.norm[2 * kMlpIn + 1] + Fs.net.norm[2 * kMlpIn]);
                const double vp = v_plain > 1.0e-30 ? v_plain : 1.0e-30;
                if (fabs(r) >= (double)Fs.gate && r < 60.0) v0 = vp * exp(r);
            }
            res = newton(v0, Fs.k_fallback, its, true);
            if (res == 0) {
                fb = 1;
                res = newton(v_plain, 20, its2, false);
            }
        } else {
            res = newton(v_plain, 20, its, false);
        }
        if constexpr (kLearn) {
            if (!active) return;
        }
        Fs.nstat[i] = Fs.nstat[i] + 1;
        Fs.nstat[ld + i] = Fs.nstat[ld + i] + its + its2;
        Fs.nstat[2 * ld + i] = Fs.nstat[2 * ld + i] + fb;
        Fs.nstat[3 * ld + i] = Fs.nstat[3 * ld + i] + (res == 0 ? 1 : 0);
        Fs.nstat[4 * ld + i] = Fs.nstat[4 * ld + i] + (res == 2 ? 1 : 0);
        if (Fs.harvest && Fs.harvest_every > 0 && nt % Fs.harvest_every == 0 && nt > Fs.harvest_t0) {
            // this is harvest step k (0-based) since harvest_t0
            const int64_t k = (nt - 1) / Fs.harvest_every - Fs.harvest_t0 / Fs.harvest_every;
            const int64_t row = k * nf + i;
            if (row < Fs.harvest_cap) {
                float *hr = Fs.harvest + row * 7;
                float feat[kMlpIn];
                features(feat);
                for (int k = 0; k < kMlpIn; k++) hr[k] = feat[k];
                hr[6] = (float)log(v_trial > 1.0e-30 ? v_trial : 1.0e-30);
            }
        }
        if (C.TPV == 105 && v_trial < FR(46)) v_trial = FR(46);
        FR(20) = stateTmp;
        FR(23) = thetaPcTmp;
        for (int j = 1; j <= 2; j++)
            T[j] = taoc_old * 0.5 * (rate[j] / rate[3]) + taoc_new * 0.5 * (trial[j] / trial[3]);
        FR(78) = T[0];
        FR(79) = T[1];
        FR(80) = T[2];
        FR(47) = v_trial;
        FR(48) = sqrt(T[1] * T[1] + T[2] * T[2]);
        if (Fs.tph) {  // onFaultTPHist, src/faulting.f90:301-302
            Fs.tph[i * Fs.nstep + (nt - 1)] = FR(47);
            Fs.tph[(nf + i) * Fs.nstep + (nt - 1)] = FR(48);
        }
        double acc[3], xyzAcc[3], xyzR[3];
        acc[0] = -rate[0] / dt - slip[0] / dt / dt;
        acc[1] = (v_trial * (trial[1] / trial[3]) - rate[1]) / dt;
        acc[2] = (v_trial * (trial[2] / trial[3]) - rate[2]) / dt;
        for (int j = 0; j < 3; j++) {
            xyzAcc[j] = acc[0] * UN(j) + acc[1] * US(j) + acc[2] * UD(j);
            xyzR[j] = Fsl[j] + Fma[j];
        }
        for (int j = 0; j < 3; j++) {
            Fsl[j] = (-xyzAcc[j] + xyzR[j] / mm) * mr;
            Fma[j] = (xyzAcc[j] + xyzR[j] / ms) * mr;
            FR(31 + j) = PV(1, nm, j) + (xyzAcc[j] + xyzR[j] / ms) * dt;
            FR(34 + j) = PV(0, ns, j) + (-xyzAcc[j] + xyzR[j] / mm) * dt;
        }
    }
    if constexpr (kLearn) {
        if (!active) return;
    }
    // storeOnFaultStationQuantSCEC, src/faulting.f90:518-541
    const int sta = Fs.station[i];
    if (sta >= 0) {
        double *h = Fs.hist + ((int64_t)sta * Fs.nstep + (nt - 1)) * 12;
        h[0] = time;
        h[1] = rate[1];
        h[2] = rate[2];
        h[3] = FR(20);
        h[4] = slip[1];
        h[5] = slip[2];
        h[6] = slip[0];
        h[7] = T[1];
        h[8] = T[2];
        h[9] = T[0];
        h[10] = FR(51) + FR(42);
        h[11] = FR(52);
    }
    // showSourceDynamics, src/faulting.f90:343-365: the hypocentre pair's
    // state after the friction solve, kept per step for the host to print
    if (Fs.hyp) {
        const int hy = Fs.hyp[i];
        if (hy >= 0) {
            double *o = Fs.hyp_hist + ((int64_t)(nt - 1) * Fs.nhyp + hy) * kHypVals;
            constexpr int slots[kHypVals - 1] = {78, 79, 80, 23, 73, 71, 72, 74, 75, 76, 77, 1, 2, 3, 9, 10, 20, 81};
            o[0] = time;
#pragma unroll
            for (int k = 0; k < kHypVals - 1; k++) o[k + 1] = FR(slots[k]);
        }
    }
    // storeRuptureTime, src/faulting.f90:329-341
    if (Fs.fnft[i] > 5000.0 && rate[3] >= C.slipRateThres) Fs.fnft[i] = time;
    // source-evolution record (output_src_evol, src/library_output.f90:297-312)
    if (M.rec) {
        const Recorder rc = *M.rec;
        const int32_t r = (nt - 1) / rc.every, lo = M.clk->rec_lo;
        if ((nt - 1) % rc.every == 0 && r >= lo && r - lo < rc.capacity)
            rc.slip_rate[(int64_t)(r % rc.capacity) * nf + i] = FR(47);
    }
#undef FR
    // F/m (src/driver.f90:29) and the pair's update for the next step
    const double msl = N.mass[ns], mma = N.mass[nm];
    if (mode == 1) {
        for (int k = 0; k < 3; k++) {
            N.acc[blk_at<3>(ns, k)] = Fsl[k] / msl;
            N.acc[blk_at<3>(nm, k)] = Fma[k] / mma;
        }
        return;
    }
    double vs[3], vm[3];
    for (int k = 0; k < 3; k++) {
        const double vs0 = PV(0, ns, k), vm0 = PV(1, nm, k);
        const double us0 = PU(0, ns, k), um0 = PU(1, nm, k);
        vs[k] = vs0 + Fsl[k] / msl * dt;
        vm[k] = vm0 + Fma[k] / mma * dt;
        const double us1 = us0 + vs[k] * dt, um1 = um0 + vm[k] * dt;
        if constexpr (sizeof(R) == sizeof(float)) {
            Fs.sv[k * ld + i] = vs[k];
            Fs.sv[(3 + k) * ld + i] = vm[k];
            Fs.su[k * ld + i] = us1;
            Fs.su[(3 + k) * ld + i] = um1;
        }
        node_store<R>(N, ns, k, vs[k], us1);
        node_store<R>(N, nm, k, vm[k], um1);
    }
    flag_nan(N, ns, nt + 1, vs[0], vs[1], vs[2]);
    flag_nan(N, nm, nt + 1, vm[0], vm[1], vm[2]);
}

// The start of a step as a launch of its own, one block per scenario (used
// when a step has no regular-element launch to carry it).
template <typename R>
__global__ void __launch_bounds__(64) step_start_kernel(const Member<R> *__restrict__ Ms)
{
    step_start_block<R>(Ms[blockIdx.x]);
}

// thermop, src/thermop.f90:1-40: pore pressure change fric(51) and
// temperature fric(52) of every pair from its history of slip rate and shear
// traction up to the previous step.  One wavefront per pair: lanes take
// every 64th past step (the same terms as the reference, each in its
// operation order), then the lane sums are added in a fixed butterfly order,
// so the result is deterministic and differs from the reference's sequential
// sum only by rounding.
template <typename R>
__global__ void __launch_bounds__(256) thermop_kernel(const Member<R> *__restrict__ Ms,
                                                      const uint16_t *__restrict__ bmem)
{
    const unsigned gb = blockIdx.x;
    const Member<R> &M = Ms[member_of(bmem, gb)];
    const int64_t tid = (int64_t)(gb - (unsigned)M.blk0[kKTp]) * blockDim.x + threadIdx.x;
    const int64_t i = tid >> 6;  // whole wavefronts share i (blocks are multiples of 64)
    const int lane = (int)(threadIdx.x & 63);
    const FaultState Fs = M.F;
    if (i >= Fs.nf) return;
    const StepConsts C = M.C;
    const int nt = M.clk->nt;
    const int64_t ld = Fs.ld, nf = Fs.nf;
    double *fr = Fs.fric;
#define FR(k) fr[(int64_t)((k) - 1) * ld + i]
    const double dt = C.dt, hh = C.fric_tp_h * C.fric_tp_h;
    const double omega = FR(16), kapa = FR(17);
    const double *hv = Fs.tph + i * Fs.nstep, *ht = Fs.tph + (nf + i) * Fs.nstep;
    double sp = 0.0, st = 0.0;
    for (int j = 1 + lane; j <= nt - 1; j += 64) {
        const double w = fabs(ht[j - 1]) * hv[j - 1];
        const double m = (double)(nt - j);
        double ker = -kapa / (omega - kapa) / sqrt(4.0 * kapa * m * dt + 2.0 * hh);
        ker = ker + omega / (omega - kapa) / sqrt(4.0 * omega * m * dt + 2.0 * hh);
        sp = sp + w * ker * dt;
        const double kt = 1.0 / sqrt(4.0 * kapa * m * dt + 2.0 * hh);
        st = st + w * kt * dt;
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        sp = sp + __shfl_xor(sp, off, 64);
        st = st + __shfl_xor(st, off, 64);
    }
    if (lane == 0) {
        const double gama = FR(19) / FR(18);
        FR(51) = sp * gama / sqrt(C.pi);
        FR(52) = st / FR(18) / sqrt(C.pi) + FR(41);
    }
#undef FR
}

// Ground-motion record (output_gm, src/library_output.f90:267-279): the
// listed nodes' velocities at a record step, taken where the reference's
// velArr holds them (after this step's update, before its forces).
template <typename R>
__global__ void __launch_bounds__(256) record_kernel(const Member<R> *__restrict__ Ms, const uint16_t *__restrict__ bmem)
{
    const unsigned gb = blockIdx.x;
    const Member<R> &M = Ms[member_of(bmem, gb)];
    const int64_t t = (int64_t)(gb - (unsigned)M.blk0[kKRecord]) * blockDim.x + threadIdx.x;
    const Recorder rc = *M.rec;
    if (t >= rc.n_nodes) return;
    const int nt = M.clk->nt;
    if ((nt - 1) % rc.every != 0) return;
    const int32_t r = (nt - 1) / rc.every, lo = M.clk->rec_lo;
    if (r < lo || r - lo >= rc.capacity) return;
    const int64_t slot = r % rc.capacity;
    const int node = rc.nodes[t];
    const int64_t nn = M.N.nn;
#pragma unroll
    for (int k = 0; k < 3; k++) rc.vel[(slot * rc.n_nodes + t) * 3 + k] = exact_v(M, node, k);
}

// ------------------------------------------------------------------------
// launchers (host side)
static inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

// One launch; with timer events t0 / t1 the dispatch itself records its start
// and end (hipExtLaunchKernelGGL), so timing adds no packets to the stream.
#define EQD_GO(KERNEL, G, B, L, S, ...)                                                           \
    do {                                                                                          \
        auto kfn_ = KERNEL;                                                                       \
        if (t0 || t1)                                                                             \
            hipExtLaunchKernelGGL(kfn_, dim3(G), dim3(B), (uint32_t)(L), S, t0, t1, 0, __VA_ARGS__); \
        else                                                                                      \
            kfn_<<<G, B, L, S>>>(__VA_ARGS__);                                                    \
    } while (0)

template <typename R>
hipError_t launch_elem(const Member<R> *M, const LaunchGrid &g, int q, int dp, int hg, int mass, int ops,
                       int nstart, hipStream_t s, hipEvent_t t0, hipEvent_t t1)
{
    if (nstart < 0) return hipErrorInvalidValue;
    if (g.grid == 0 && nstart == 0) return hipSuccess;
    // leading blocks for the step start of nstart scenarios (see elem_regular_kernel)
    const int grid = g.grid + step_start_skip(nstart);
    if (g.block < 64 || g.block > 256 || g.block % 64) return hipErrorInvalidValue;
    if (ops < kOpsRows || ops > kOpsGeom) return hipErrorInvalidValue;
    // geometry rows exist in the FP32 bulk mode only
    if (ops == kOpsGeom && sizeof(R) != sizeof(float)) return hipErrorInvalidValue;
    constexpr int kGeoOr = sizeof(R) == sizeof(float) ? kOpsGeom : kOpsCompact;
    // class tables in LDS (at most 16 KiB) or operators from memory
    const size_t lds = ops == kOpsLds ? g.lds : 0;
#define EQD_LAUNCH(Q, DP, HG, MASS)                                                                               \
    do {                                                                                                          \
        if (ops == kOpsLds)                                                                                       \
            EQD_GO((elem_regular_kernel<R, Q, DP, HG, MASS, kOpsLds>), grid, g.block, lds, s, M, g.bmem, g.remap,  \
                   nstart);                                                                                       \
        else if (ops == kOpsGeom)                                                                                 \
            EQD_GO((elem_regular_kernel<R, Q, DP, HG, MASS, kGeoOr>), grid, g.block, 0, s, M, g.bmem, g.remap,     \
                   nstart);                                                                                       \
        else if (ops == kOpsCompact)                                                                              \
            EQD_GO((elem_regular_kernel<R, Q, DP, HG, MASS, kOpsCompact>), grid, g.block, 0, s, M, g.bmem,         \
                   g.remap, nstart);                                                                              \
        else                                                                                                      \
            EQD_GO((elem_regular_kernel<R, Q, DP, HG, MASS, kOpsRows>), grid, g.block, 0, s, M, g.bmem, g.remap,   \
                   nstart);                                                                                       \
    } while (0)
#define EQD_HG(Q, DP, MASS)                         \
    do {                                            \
        if (hg == 1) EQD_LAUNCH(Q, DP, 1, MASS);    \
        else if (hg == 2) EQD_LAUNCH(Q, DP, 2, MASS); \
        else EQD_LAUNCH(Q, DP, 0, MASS);            \
    } while (0)
    if (q && mass) EQD_HG(true, false, true);
    else if (q) EQD_HG(true, false, false);
    else if (dp) EQD_HG(false, true, true);
    else if (mass) EQD_HG(false, false, true);
    else EQD_HG(false, false, false);
#undef EQD_HG
#undef EQD_LAUNCH
    return hipGetLastError();
}

template <typename R>
hipError_t launch_pml(const Member<R> *M, const LaunchGrid &g, int hg, int mass, int ops, hipStream_t s,
                      hipEvent_t t0, hipEvent_t t1)
{
    if (g.grid == 0) return hipSuccess;
    if (g.block < 64 || g.block > 256 || g.block % 64) return hipErrorInvalidValue;
    if (ops < kOpsRows || ops > kOpsGeom) return hipErrorInvalidValue;
    if (ops == kOpsGeom && sizeof(R) != sizeof(float)) return hipErrorInvalidValue;
    constexpr int kGeoOr = sizeof(R) == sizeof(float) ? kOpsGeom : kOpsCompact;
    // class tables, or each thread's staged dN/dx (24 values); geometry rows
    // keep dN/dx in registers
    const size_t lds = ops == kOpsLds ? g.lds
                                      : (ops == kOpsGeom ? 0 : (EQD_STAGE_SHP_PML ? (size_t)g.block * 24 * sizeof(R) : 0));
#define EQD_LAUNCH(HG, MASS)                                                                              \
    do {                                                                                                  \
        if (ops == kOpsLds)                                                                               \
            EQD_GO((elem_pml_kernel<R, HG, MASS, kOpsLds>), g.grid, g.block, lds, s, M, g.bmem, g.remap);     \
        else if (ops == kOpsGeom)                                                                         \
            EQD_GO((elem_pml_kernel<R, HG, MASS, kGeoOr>), g.grid, g.block, lds, s, M, g.bmem, g.remap);  \
        else if (ops == kOpsCompact)                                                                      \
            EQD_GO((elem_pml_kernel<R, HG, MASS, kOpsCompact>), g.grid, g.block, lds, s, M, g.bmem, g.remap); \
        else                                                                                              \
            EQD_GO((elem_pml_kernel<R, HG, MASS, kOpsRows>), g.grid, g.block, lds, s, M, g.bmem, g.remap);    \
    } while (0)
    if (mass) {
        if (hg == 1) EQD_LAUNCH(1, true);
        else if (hg == 2) EQD_LAUNCH(2, true);
        else EQD_LAUNCH(0, true);
    } else {
        if (hg == 1) EQD_LAUNCH(1, false);
        else if (hg == 2) EQD_LAUNCH(2, false);
        else EQD_LAUNCH(0, false);
    }
#undef EQD_LAUNCH
    return hipGetLastError();
}

template <typename R>
hipError_t launch_node3(const Member<R> *M, const LaunchGrid &g, hipStream_t s, hipEvent_t t0, hipEvent_t t1)
{
    if (g.grid == 0) return hipSuccess;
    if (g.block < 64 || g.block > 256 || g.block % 64) return hipErrorInvalidValue;
    EQD_GO(node3_kernel<R>, g.grid, g.block, 0, s, M, g.bmem, g.remap);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_node_pml(const Member<R> *M, const LaunchGrid &g, hipStream_t s, hipEvent_t t0, hipEvent_t t1)
{
    if (g.grid == 0) return hipSuccess;
    if (g.block < 64 || g.block > 256 || g.block % 64) return hipErrorInvalidValue;
    EQD_GO(node_pml_kernel<R>, g.grid, g.block, 0, s, M, g.bmem, g.remap);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_pending(const Member<R> *M, const LaunchGrid &g3, const LaunchGrid &gp, hipStream_t s)
{
    if (g3.grid > 0) update3_from_acc_kernel<R><<<g3.grid, g3.block, 0, s>>>(M, g3.bmem);
    if (gp.grid > 0) update_pml_from_acc_kernel<R><<<gp.grid, gp.block, 0, s>>>(M, gp.bmem);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_fault(const Member<R> *M, const LaunchGrid &g, int learn, hipStream_t s, hipEvent_t t0,
                        hipEvent_t t1)
{
    if (g.grid == 0) return hipSuccess;
    // the learned variant evaluates the network one wavefront at a time
    if (g.block != 64 && g.block != 128) return hipErrorInvalidValue;
    if (learn)
        EQD_GO((fault_kernel<R, true>), g.grid, g.block, 0, s, M, g.bmem);
    else
        EQD_GO((fault_kernel<R, false>), g.grid, g.block, 0, s, M, g.bmem);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_record(const Member<R> *M, const LaunchGrid &g, hipStream_t s, hipEvent_t t0, hipEvent_t t1)
{
    if (g.grid == 0) return hipSuccess;
    if (g.block < 64 || g.block > 256 || g.block % 64) return hipErrorInvalidValue;
    EQD_GO(record_kernel<R>, g.grid, g.block, 0, s, M, g.bmem);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_thermop(const Member<R> *M, const LaunchGrid &g, hipStream_t s, hipEvent_t t0, hipEvent_t t1)
{
    if (g.grid == 0) return hipSuccess;
    if (g.block < 64 || g.block > 256 || g.block % 64) return hipErrorInvalidValue;
    EQD_GO(thermop_kernel<R>, g.grid, g.block, 0, s, M, g.bmem);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_step_start(const Member<R> *M, int nmem, hipStream_t s, hipEvent_t t0, hipEvent_t t1)
{
    if (nmem <= 0) return hipSuccess;
    EQD_GO(step_start_kernel<R>, nmem, 64, 0, s, M);
    return hipGetLastError();
}

#define EQD_INSTANTIATE(R)                                                                                     \
    template hipError_t launch_elem<R>(const Member<R> *, const LaunchGrid &, int, int, int, int, int, int, hipStream_t, \
                                       hipEvent_t, hipEvent_t);                                                       \
    template hipError_t launch_pml<R>(const Member<R> *, const LaunchGrid &, int, int, int, hipStream_t, hipEvent_t,  \
                                      hipEvent_t);                                                                    \
    template hipError_t launch_node3<R>(const Member<R> *, const LaunchGrid &, hipStream_t, hipEvent_t, hipEvent_t);  \
    template hipError_t launch_node_pml<R>(const Member<R> *, const LaunchGrid &, hipStream_t, hipEvent_t,            \
                                           hipEvent_t);                                                               \
    template hipError_t launch_pending<R>(const Member<R> *, const LaunchGrid &, const LaunchGrid &, hipStream_t);    \
    template hipError_t launch_fault<R>(const Member<R> *, const LaunchGrid &, int, hipStream_t, hipEvent_t,           \
                                        hipEvent_t);                                                                  \
    template hipError_t launch_step_start<R>(const Member<R> *, int, hipStream_t, hipEvent_t, hipEvent_t);           \
    template hipError_t launch_thermop<R>(const Member<R> *, const LaunchGrid &, hipStream_t, hipEvent_t, hipEvent_t);\
    template hipError_t launch_record<R>(const Member<R> *, const LaunchGrid &, hipStream_t, hipEvent_t, hipEvent_t);
EQD_INSTANTIATE(double)
EQD_INSTANTIATE(float)
#undef EQD_INSTANTIATE

}  // namespace eqd
