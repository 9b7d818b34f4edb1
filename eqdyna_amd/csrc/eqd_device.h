// Device-side data layout of the eqdyna_amd time step.
//
// Everything the step touches stays resident in HBM in blocked
// structure-of-arrays form (blk_at): value c of item t lives at
// ((t / 64) * W + c) * 64 + t % 64, so consecutive threads read consecutive
// addresses and one wavefront's W values of its 64 items form one contiguous
// region:
//  * elements are split by class (regular hex/wedge vs PML) and renumbered
//    within the class in ascending reference order;
//  * nodal arrays hold 3 values per node in the same blocks of 64 nodes;
//  * the element-force buffers hold, per element and local node, the force the
//    element contributes to that node (for PML nodes in factored form); a
//    node-to-element table (CSR, ascending element order) lets each node sum
//    its contributions without atomics.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "eqd_mlp.h"

namespace eqd {

// CSR entry: bit 31 = PML element, bits 28..30 = local node a, bits 0..27 =
// element index within its class.
constexpr uint32_t kEntPml = 0x80000000u;
constexpr uint32_t kEntIdxMask = 0x0FFFFFFFu;
constexpr int kEntLocalShift = 28;
constexpr uint32_t kEntNone = 0xFFFFFFFFu;  // an unused slot of the ELL gather table
constexpr int kEll3 = 8;
// values recorded per step for a hypocentre pair (showSourceDynamics,
// src/faulting.f90:343-365): timeElapsed, fric 78, 79, 80, 23, 73, 71, 72,
// 74, 75, 76, 77, 1, 2, 3, 9, 10, 20, 81
constexpr int kHypVals = 19;

struct StepConsts {
    double dt, w, rdampk, rdampm, gterm, ccosphi, sinphi, exp_dt_tv, kapa_hg, fric_tp_h, pi;
    double slipRateThres, xsource, ysource, zsource, nucR, nucRuptVel, nucdtau0, nucT;
    int32_t C_elastic, C_Q, C_hg, C_nuclea, friclaw, TPV, insertFaultType, nucfault;
};

// Node state.  FP64 mode: vel and disp in FP64.  FP32 bulk mode: vel32 in
// FP32 and the displacement as an unevaluated FP32 pair dhi + dlo (dhi the
// displacement rounded to FP32, dlo what that rounding dropped), about 48
// significant bits, so the accumulation u += v dt keeps FP64-like accuracy
// while the element kernels gather 4-byte v and dhi; the split nodes of the
// fault keep their own FP64 state (FaultState::sv, su), of which vel32, dhi
// and dlo hold the rounded copy.
struct NodeState {
    int64_t nn;
    double *vel;   // 3 per node (blk_at<3>): velArr (== v1 for 3-dof nodes); FP64 mode
    double *disp;  // 3 per node (blk_at<3>): dispArr; FP64 mode
    float *vel32;  // 3 per node (blk_at<3>); FP32 bulk mode
    // displacement in the FP32 bulk mode, 3 per node (blk_at<3>): dhi the FP32
    // value the element kernels gather, dlo the residual u - dhi as bfloat16
    // bits, so the node kernels carry u to about 32 significant bits
    float *dhi;
    uint16_t *dlo;
    const int32_t *pair_slot;  // [nn] FP32 bulk mode: 2 * pair + (0 slave, 1 master) of a split node, else -1
    double *acc;   // 3 per node (blk_at<3>): pending F/m of 3-dof nodes
    const double *mass;    // [nn] nodal mass (nodalMassArr of the node's equations)
    const float *mass32;   // [nn] the same rounded to FP32, read by the FP32 bulk mode's node kernels
    const int32_t *rowptr; // [nn+1]
    const uint32_t *ent;   // CSR entries
    unsigned long long *nanflag; // first NaN velocity: step << 32 | node + 1 (all ones: none)
};

// layout of one operator row: dN_a/dx_k at kOpShp + a*3+k, det, lambda, mu,
// Vp, rho (viscous hourglass only), SS (6); kOpB-wide per-element rows add
// Phi_a of mode m at kOpW + m*8+a, kOpBc-wide rows the compact hourglass
// operators instead (phi_compact)
constexpr int kOpShp = 0, kOpDet = 24, kOpLam = 25, kOpMu = 26, kOpVp = 27, kOpRho = 28, kOpSs = 29, kOpW = 35,
              kOpB = 67, kOpBc = kOpW + 16;
// Geometry rows (kOpsGeom, FP32 bulk mode on meshes whose elements all
// differ): the coordinates of local nodes 1..7 relative to node 0 at
// (a - 1) * 3 + j, then lambda, mu, Vp, rho at kOpG - 4 ..; the kernels
// rebuild dN/dx, det, SS and the compact hourglass operators from them.
constexpr int kOpG = 25, kGeoMat = 21;

// value k of element t's per-element operator row of W values (AoSoA, blocks
// of 64 elements), loaded non-temporal: each value is read once per step, and
// keeping it out of the caches leaves them to the node state the element
// kernels gather
template <typename R, int W = kOpB>
__device__ __forceinline__ R opb_at(const R *opb, int64_t t, int k)
{
    return __builtin_nontemporal_load(&opb[((t >> 6) * W + k) * 64 + (t & 63)]);
}

// Value s of item t (element or PML node) of a per-item array of W values:
// AoSoA in blocks of 64 items, so a wavefront's accesses of one item row are
// one contiguous region of W * 64 values instead of W streams far apart.
// Arrays in this layout hold (n + 63) / 64 * 64 * W values.
template <int W>
__host__ __device__ __forceinline__ int64_t blk_at(int64_t t, int s)
{
    return ((t >> 6) * W + s) * 64 + (t & 63);
}

// The FP64 PML elements' node slots and summed stresses are read in pairs:
// values 2j and 2j+1 of item t sit side by side, so one lane loads or stores
// both at once (16 bytes) and a wavefront's access to a pair is one
// contiguous run of 128 values (W even; the same footprint as blk_at<W>).
// c4 FP64 PML element kernel 0.288 -> 0.252 ms, step 1.006 -> 0.966 ms; in
// FP32 (8-byte pairs) the same layout measured no faster on c4 and 1.4 %
// slower on the 64-scenario ensemble, so the FP32 mode keeps blk_at.
// The regular elements' forces on their nodes, component c of local node a
// of element t: FP64 keeps x and y of a node as a 16-byte pair (rows 2a,
// 2a+1 of a 64-element block interleaved) and z in rows 16 + a, so a node's
// force is one paired and one single access; FP32 keeps blk_at<24> slot
// a*3+c.  Both fill 24 rows of 64 values per block.
template <typename R>
__host__ __device__ __forceinline__ int64_t reg_ef_at(int64_t t, int a, int c)
{
    if constexpr (sizeof(R) == sizeof(double)) {
        if (c < 2) return ((t >> 6) * 24 + 2 * a) * 64 + (t & 63) * 2 + c;
        return ((t >> 6) * 24 + 16 + a) * 64 + (t & 63);
    } else {
        return blk_at<24>(t, a * 3 + c);
    }
}

template <typename R, int W>
__host__ __device__ __forceinline__ int64_t pml_at(int64_t t, int s)
{
    if constexpr (sizeof(R) == sizeof(double)) return ((t >> 6) * W + (s & ~1)) * 64 + (t & 63) * 2 + (s & 1);
    else return blk_at<W>(t, s);
}
constexpr int kEfReg = 24, kEfPml = 48;

// Where an element kernel finds its operators: class tables in global memory
// or per-element rows with Phi; class tables staged in LDS; per-element rows
// with the compact hourglass operators; per-element geometry rows.
constexpr int kOpsRows = 0, kOpsLds = 1, kOpsCompact = 2, kOpsGeom = 3;

// Hourglass base vectors Gamma_m (ha of calcSSPhi4Hrgls,
// src/assembleGlobalMass.f90:332-335): bit a set where Gamma_ma = -1.
constexpr unsigned kGammaNeg[4] = {0x3Cu, 0x96u, 0xAAu, 0xA5u};

// The reference's Phi_ma = (Gamma_ma - h_m . dN_a/dx) / s_m
// (src/assembleGlobalMass.f90:351-372), with
// h_m = sum_b Gamma_mb x_b and s_m the rms of the numerator over a, so four
// numbers per mode carry its eight values: c[m] = 1/s_m and c[4 + 3m + k] =
// h_mk / s_m.  Rebuilt here with no fused operations, in this order, on the
// host and on the device alike.
template <typename R>
__host__ __device__ __forceinline__ R phi_compact(const R *c, int m, int a, R s0, R s1, R s2)
{
    const R r = ((kGammaNeg[m] >> a) & 1u) ? -c[m] : c[m];
    R h = c[4 + 3 * m] * s0;
    h = h + c[5 + 3 * m] * s1;
    h = h + c[6 + 3 * m] * s2;
    return r - h;
}

template <typename R>
struct RegElems {
    int64_t n;
    const int32_t *ien;   // 8 per element (blk_at<8>), 0-based node ids
    R *stress;            // 6, or 12 with Q, per element (blk_at)
    const R *emass;       // 24 per element (blk_at<24>; inertia/gravity term only)
    const double *porep;  // [n]      (plastic only)
    double *pstrain;      // [n]      (plastic only)
    const uint8_t *qidx;  // [n]      (Q only) row of qtab
    const double *qtab;   // [16][6]  exp(-dt/tau), 1-exp(-dt/tau), 8*w_p, 8*w_s, c_p, c_s
    R *ef;                // force on local node a, component c at slot a*3+c (blk_at<kEfReg>)
    // Operators: with classes (optab set), elements whose operators are
    // bit-identical share one row of optab (kOpW values, at opc[t]) and of
    // phitab (32 values, at phic[t]); otherwise each element has its own
    // row in opb (opb_at): kOpBc values with the compact hourglass operators
    // when opb_compact is set, else kOpB values with Phi.
    const uint32_t *opc;
    const R *optab;
    const uint32_t *phic;
    const R *phitab;
    int32_t n_opc, n_phic;  // rows of optab and phitab
    const R *opb;
    int32_t opb_compact;
    int32_t opb_geom;  // opb holds geometry rows (kOpG values)
};

template <typename R>
struct PmlElems {
    int64_t n;
    const int32_t *ien;    // 8 per element (blk_at<8>)
    R *stress;             // 21 per element (blk_at<21>): 15 split + 6 initial
    const R *emass;        // 24 per element (blk_at<24>)
    const R *damp;         // 3 per element (blk_at<3>): element-centre damping d_x, d_y, d_z
    const uint8_t *mask12; // [n] bit a: local node a carries 12 components
    R *ef;                 // 6 slots per local node a (blk_at<kEfPml>, slot a*6+k): a 3-component
                           // node's force at k = 0..2; a 12-component node's det*w*dN_a/dx at
                           // k = 0..2 (its 9 split forces follow from ssum) and force
                           // components 10-12 at k = 3..5
    R *ssum;               // 6 per element (blk_at<6>): summed stresses xx, yy, zz, xy, xz, yz
    // operators as in RegElems
    const uint32_t *opc;
    const R *optab;
    const uint32_t *phic;
    const R *phitab;
    int32_t n_opc, n_phic;
    const R *opb;
    int32_t opb_compact;
    int32_t opb_geom;
};

struct PmlNodes {
    int64_t n;
    const int32_t *node;  // [n]
    const double *damp;   // 3 per node (blk_at<3>): d_x, d_y, d_z at the node
    const float *damp32;  // the same rounded to FP32, read by the FP32 bulk mode
    double *v12;          // 12 per node (blk_at<12>): split velocities (v1 of the 12 equations); FP64 mode
    float *v12f;          // the same in FP32; FP32 bulk mode
    double *acc;          // 12 per node (blk_at<12>): pending F/m
};

struct FaultState {
    int64_t nf;
    int64_t ld;             // row stride of fric and nstat: nf rounded up to the fault kernel's block
    int32_t *nstat;         // [5][ld] Newton counters: solves, iterations, fallbacks, unconverged,
                            // stopped at the roundoff floor
    // learned Newton guess (when the kernel is launched in learned mode)
    MlpNet net;
    int32_t k_fallback;     // 0: the reference's solve; > 0: the faster mode's iteration policy
    float gate;             // predicted |ln(V/V_prev)| below which the plain guess is kept
    // harvest of (inputs, converged ln V): rows [harvest_cap][7], nf rows at
    // every step nt > harvest_t0 with nt % harvest_every == 0, in step order
    // (no harvest when harvest is null)
    float *harvest;
    int64_t harvest_cap;
    int32_t harvest_every, harvest_t0;
    const int32_t *nsmp;    // [2][nf] slave, master (0-based node ids)
    const double *un, *us, *ud; // [3][nf]
    const double *arn;      // [nf]
    const double *fms, *fmm; // [nf] fnms of slave and master
    const double *xyz;      // [3][nf] slave coordinates
    double *sv, *su;        // FP32 bulk mode: FP64 velocity and displacement of the pair's nodes,
                            // [(side * 3 + k) * ld + pair], side 0 slave, 1 master (else null)
    double *fric;           // [100][ld]
    double *fnft;           // [nf]
    const int32_t *station; // [nf] on-fault station (0-based) or -1
    double *hist;           // [non][nstep][12]
    // hypocentre pairs (showSourceDynamics): row of the pair or -1, and per
    // step and row the kHypVals values the reference prints
    const int32_t *hyp;     // [nf] or null
    double *hyp_hist;       // [nstep][nhyp][kHypVals]
    int32_t nhyp;
    int32_t nstep;
    // thermal pressurization (friclaw 5): slip rate and shear traction of
    // every pair at every step, [2][nf][nstep] (onFaultTPHist), else null
    double *tph;
};

struct OffStations {
    int32_t nrow;          // 6 * noff
    const int32_t *idh;    // [3][nrow]: node (0-based), component (0-based), kind (1 disp, 2 vel)
    double *hist;          // [nstep][nrow+1]
};

// The step clock of one scenario, in device memory.  The clock kernel
// advances it at the start of every step (nt += 1, time += dt, the same FP64
// additions as the reference's timeElapsed), so a captured step replays
// unchanged; every kernel of the step reads nt, time and mode from here.
struct Clock {
    double time;      // timeElapsed after this step's advance
    int32_t nt;       // current step
    int32_t mode;     // 1 on the last step of a run: store F/m instead of the fused update
    int32_t run_end;  // the step at which the current run ends
    int32_t rec_lo;   // first record index the host can take in this run (Recorder)
};

// launch kinds of one step
enum LaunchKind : int {
    kKElem = 0, kKPml, kKNode3, kKNodePml, kKFault, kKStation, kKClock, kKPend3, kKPendPml, kKTp, kKRecord, kKinds
};

// Ground-motion and source-evolution records (the data of output_gm and
// output_src_evol, src/library_output.f90:267-312): at every step nt with
// (nt - 1) % every == 0, record r = (nt - 1) / every holds the velocities of
// the listed nodes and fric(47) of every pair.  The device buffers are a ring
// of `capacity` records (slot r % capacity); a run writes only the records in
// [clk->rec_lo, clk->rec_lo + capacity), the ones whose slots the host has
// already copied out, and drops the rest.
struct Recorder {
    int32_t every, capacity, n_nodes, pad;
    const int32_t *nodes;  // [n_nodes] 0-based node ids
    double *vel;           // [capacity][n_nodes][3]
    double *slip_rate;     // [capacity][nf]
};

// Everything one scenario's step reads: every kernel takes an array of these
// (one per scenario of a batched launch) and finds its scenario from the
// launch's block-to-scenario table; blk0[kind] is the scenario's first block
// in that kind's launch.
template <typename R>
struct Member {
    RegElems<R> E;
    PmlElems<R> P;
    NodeState N;
    PmlNodes PN;
    FaultState F;
    OffStations S;
    StepConsts C;
    const int32_t *node3_free;  // 3-component nodes that are not split nodes
    int64_t n_node3_free;
    // the free nodes' gather entries slot-major by position in node3_free
    // (blk_at<kEll3>: entry k of free node i at slot k), unused slots
    // kEntNone; null when a free node has more than kEll3 entries (then the
    // node kernel reads the CSR table)
    const uint32_t *ell3;
    const int32_t *node3_all;   // all 3-component nodes
    int64_t n_node3_all;
    Clock *clk;
    const Recorder *rec;  // null when nothing is recorded
    int32_t blk0[kKinds];
};

// one launch of one kind: grid and block size, the block-to-scenario table
// (device memory, null for a single scenario), XCD remapping, dynamic LDS
struct LaunchGrid {
    unsigned grid;
    int block;
    int remap;
    const uint16_t *bmem;
    size_t lds;
};

#ifdef __HIPCC__
// Block b of a launch runs on XCD b % 8 (blocks are dealt out round-robin).
// With remap the blocks are renumbered so that each XCD takes one contiguous
// range of the launch's work: neighbouring elements, which share nodes, are
// then served by the same L2.  A bijection of [0, gridDim.x).
// skip: leading blocks of the launch that do other work (a multiple of 8).
__device__ __forceinline__ unsigned launch_block(int remap, unsigned skip = 0)
{
    const unsigned b = blockIdx.x - skip, g = gridDim.x - skip;
    if (!remap) return b;
    const unsigned x = b & 7u, i = b >> 3, q = g >> 3, r = g & 7u;
    // XCD x runs q + 1 blocks when x < r, else q
    return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

// leading blocks of an element launch that carries the step start of
// nstart scenarios
__host__ __device__ __forceinline__ int step_start_skip(int nstart) { return (nstart + 7) & ~7; }

// scenario of logical block gb (a batched launch has a table; a single
// scenario's launch has none)
__device__ __forceinline__ int member_of(const uint16_t *bmem, unsigned gb)
{
    return bmem ? __builtin_amdgcn_readfirstlane((int)bmem[gb]) : 0;
}
#endif

}  // namespace eqd
