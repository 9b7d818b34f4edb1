/*
 * eqdyna_amd: MI355X (gfx950) explicit time step for EQdyna-style dynamic
 * rupture, behind a plain C ABI.
 *
 * The reference calls its time loop as an argument-less subroutine,
 *     call driver                      (src/eqdyna3d.f90:72, body src/driver.f90:3-36)
 * with all data in module globalvar (src/globalvar.f90:3-103).  This ABI is
 * what a Fortran iso_c_binding shim replacing driver.f90 binds: the host hands
 * over globalvar's arrays exactly as they lie in Fortran memory (column-major,
 * 1-based ids, the layout listed in src/globalvar.f90:78-101), the library
 * keeps the whole mesh resident in GPU memory, advances steps, and copies the
 * state back into the same arrays so the reference's output routines
 * (src/library_output.f90) can write them unchanged.
 *
 * Conventions: every function returns 0 on success or a negative EQD_E* code;
 * eqd_last_error() gives the message.  Handles are not re-entrant; one handle
 * drives one device stream.
 */
#ifndef EQDYNA_AMD_H
#define EQDYNA_AMD_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EQD_OK 0
#define EQD_EINVAL (-1)      /* bad argument, size or unsupported combination */
#define EQD_ENAN (-2)        /* NaN velocity (reference: src/driver.f90:147-152) */
#define EQD_EHIP (-3)        /* HIP runtime error */
#define EQD_EOOM (-4)        /* device allocation failed */
#define EQD_ESTATE (-5)      /* call out of order (e.g. run before commit) */

/* Arithmetic of the bulk (element force + hourglass) path.  The split-node
 * fault, the rate-state Newton solve and the nodal update are FP64 in every
 * mode. */
#define EQD_PREC_FP64 0
#define EQD_PREC_FP32 1

/* Scalars of module globalvar that the time step reads.  Names are the
 * reference's; counts are the allocation sizes of the arrays below. */
typedef struct {
    double dt, rdampk, rdampm, w, grav, roumax, rhow, gamar;
    double ccosphi, sinphi, tv, slipRateThres;
    double xsource, ysource, zsource, nucR, nucRuptVel, nucdtau0, nucT;
    double kapa_hg, dx, R, vmaxPML, PMLb[8], pi;
    int32_t nstep;                    /* length of the station histories */
    int32_t nn, ne, neq;              /* totalNumOfNodes/Elements/Equations */
    int32_t size_eqidx;               /* sizeOfEqNumIndexArr */
    int32_t size_stress;              /* sizeOfStressDofIndexArr */
    int32_t nf;                       /* nftnd(1): split-node pairs */
    int32_t non, noff;                /* numOfOnFaultStCount, numOfOffFaultStCount */
    int32_t nPML;
    int32_t C_elastic, C_Q, C_hg, C_nuclea, friclaw, TPV, insertFaultType, nucfault;
    int32_t precision;                /* EQD_PREC_* */
    int32_t device;                   /* HIP device ordinal */
    double fric_tp_h;                 /* thermal pressurization half-width (src/thermop.f90) */
} eqd_scalars;

/* Array ids for eqd_set / eqd_get.  Shapes are Fortran (column-major). */
enum eqd_field {
    EQD_MESHCOOR = 1,            /* f64 (3, nn)          in            */
    EQD_EQNUMINDEXARR,           /* i32 (size_eqidx)     in            */
    EQD_EQNUMSTARTINDEXLOC,      /* i32 (nn)             in            */
    EQD_NUMOFDOFPERNODEARR,      /* i32 (nn)             in            */
    EQD_NODEELEMIDRELATION,      /* i32 (8, ne)          in            */
    EQD_ELEMTYPEARR,             /* i32 (ne)             in            */
    EQD_MAT,                     /* f64 (ne, 5)          in            */
    EQD_ELESHP,                  /* f64 (3, 8, ne)       in            */
    EQD_ELEDET,                  /* f64 (ne)             in            */
    EQD_ELEMASS,                 /* f64 (24, ne)         in            */
    EQD_SS,                      /* f64 (6, ne)          in            */
    EQD_PHI,                     /* f64 (8, 4, ne)       in            */
    EQD_STRESSCOMPINDEXARR,      /* i32 (ne)             in            */
    EQD_ELEPOREP,                /* f64 (ne)             in            */
    EQD_NODALMASSARR,            /* f64 (neq)            in            */
    EQD_FNMS,                    /* f64 (nn)             in            */
    EQD_NSMP,                    /* i32 (2, nf)          in            */
    EQD_UN,                      /* f64 (3, nf)          in            */
    EQD_US,                      /* f64 (3, nf)          in            */
    EQD_UD,                      /* f64 (3, nf)          in            */
    EQD_ARN,                     /* f64 (nf)             in            */
    EQD_ANONFS,                  /* i32 (3, non)         in            */
    EQD_IDHIST,                  /* i32 (3, 6*noff)      in            */
    EQD_STRESSARR,               /* f64 (size_stress)    in/out        */
    EQD_PSTRAIN,                 /* f64 (ne)             in/out        */
    EQD_V1,                      /* f64 (neq)            in/out        */
    EQD_VELARR,                  /* f64 (3, nn)          in/out        */
    EQD_DISPARR,                 /* f64 (3, nn)          in/out        */
    EQD_NODALFORCEARR,           /* f64 (neq)            in/out (holds F/m between steps) */
    EQD_FRIC,                    /* f64 (100, nf)        in/out        */
    EQD_FNFT,                    /* f64 (nf)             in/out        */
    EQD_ONFAULTHIST,             /* f64 (12, nstep, non) out: onFaultQuantHistSCECForm */
    EQD_OFFFAULTHIST,            /* f64 (6*noff+1, nstep) out: OffFaultStGramSCEC */
    EQD_FIELD_COUNT
};

typedef struct eqd_handle eqd_handle;

/* Replaces the state set up by allocInit/meshgen/assembleGlobalMass
 * (src/eqdyna3d.f90:54-70): stores the scalars; arrays follow via eqd_set. */
int eqd_create(const eqd_scalars *s, eqd_handle **h);

/* Hand over one globalvar array.  The library reads it during eqd_commit and
 * keeps no reference afterwards, so the buffer must stay valid (and unchanged)
 * until eqd_commit returns; no copy is made before that. */
int eqd_set(eqd_handle *h, int32_t field, const void *host, int64_t count);

/* Validate, build the device layout (class-sorted SoA element arrays, the
 * node-to-element gather table, PML damping, Q tables) and upload. */
int eqd_commit(eqd_handle *h);

/* Advance nsteps steps of src/driver.f90:9-34 (velDispUpdate,
 * storeOffFaultStData, assembleGlobalKU, hrglss, faulting, F/m), continuing the
 * step counter nt and timeElapsed across calls.  Returns EQD_ENAN if a velocity
 * becomes NaN; eqd_nan_info then names the node and step. */
int eqd_run(eqd_handle *h, int32_t nsteps);

/* eqd_run in two halves, so the host can work while the device steps (the
 * reference's loop is synchronous, src/driver.f90:9-34; these replace the
 * `call driver` of src/eqdyna3d.f90:72 in a host that overlaps its output
 * writing with the next chunk of steps).  eqd_run_async enqueues nsteps steps
 * and returns; eqd_wait waits for them and returns what eqd_run would have
 * (EQD_ENAN ...).  While a run is in flight eqd_record_get hands over the
 * records of the runs before it without waiting for it; every other call on
 * the handle first waits for the run (and returns its error, if any). */
int eqd_run_async(eqd_handle *h, int32_t nsteps);
int eqd_wait(eqd_handle *h);

/* Copy one array back into reference layout. */
int eqd_get(eqd_handle *h, int32_t field, void *host, int64_t count);

/* Clock of the time loop (nt, timeElapsed of src/driver.f90:9-11). */
int eqd_clock(eqd_handle *h, int32_t *nt, double *time_elapsed);

/* Node id (1-based) and step of the first NaN velocity, 0 if none. */
int eqd_nan_info(eqd_handle *h, int32_t *node, int32_t *nt);

/* Average device time per launch of each kernel over the last eqd_run,
 * measured with HIP events on the handle's stream (milliseconds), and the
 * number of launches.  names: "element", "element_pml", "fault", "node",
 * "node_pml", "station" (the step start: clock and off-fault stations; it
 * has launches of its own only in steps without regular elements, otherwise
 * it runs in the first blocks of the "element" launch). */
int eqd_kernel_time(eqd_handle *h, const char *name, double *avg_ms, int64_t *launches);

/* Per-launch event timing used by eqd_kernel_time: 0 off, k >= 1 on every
 * k-th step of each run (1: every step).  The events come from the
 * dispatches themselves and cost device time at every kernel boundary they
 * mark, so sampling keeps a timed run close to an untimed one. */
int eqd_set_timing(eqd_handle *h, int32_t on);

/* Device bytes held by the handle. */
int64_t eqd_device_bytes(eqd_handle *h);

/* Facts about the device layout built by eqd_commit: "regular_elements",
 * "pml_elements", and the number of operator classes (elements whose stored
 * operators are bit-identical share one table row; 0 = per-element arrays):
 * "op_classes_regular", "phi_classes_regular", "op_classes_pml",
 * "phi_classes_pml"; "phi_compact_regular", "phi_compact_pml": 1 when the
 * hourglass Phi is carried as four numbers per mode (it has the form of
 * src/assembleGlobalMass.f90:351-372 for every element of the class; then
 * Phi agrees with the given one to 1e-11), 0 when the given Phi is kept;
 * "geom_ops_regular", "geom_ops_pml": 1 when the FP32 bulk mode rebuilds the
 * class's operators from the elements' corner coordinates in the kernels
 * (meshes whose elements all differ; the given operators must be those of
 * the coordinates, EQD_GEOM_OPS=0 in the environment keeps them stored);
 * "node3_slot_major": 1 when the 3-component nodes gather their element
 * forces from a slot-major table of 8 entries per node (every such node has
 * at most 8 elements; EQD_NODE_ELL=0 keeps the CSR table).
 * Returns EQD_EINVAL for an unknown key. */
int eqd_layout_info(eqd_handle *h, const char *key, int64_t *value);

const char *eqd_last_error(eqd_handle *h);
int eqd_destroy(eqd_handle *h);

/* ---------------------------------------------------------------------------
 * Learned initial guess of the rate-and-state Newton solve.  No counterpart in
 * the reference, whose NewtonRaphson (src/faulting.f90:459-516) starts from
 * the previous slip rate; the solve itself and its convergence test are
 * unchanged, the network only proposes the starting point.
 *
 * Network: 6 -> 64 -> 64 -> 64 -> 1, GELU; bfloat16 weights and activations
 * on the matrix cores, float32 accumulation.  Inputs per fault node (from
 * solveRSF, src/faulting.f90:190-327): trial shear traction (Pa), effective
 * normal stress (Pa), state variable (psi of the slip law, ln theta of the
 * ageing law), a, b, ln |V_prev| (m/s); output r = ln(V / V_prev), the
 * proposal being V_prev * exp(r).
 * Parameter vector (float32, EQD_MLP_NPARAMS values): W1[64][6] b1[64]
 * W2[64][64] b2[64] W3[64][64] b3[64] w4[64] b4 in_shift[6] in_scale[6]
 * out_shift out_scale; normalised input (x - in_shift) * in_scale; r =
 * y * out_scale + out_shift.  The first EQD_MLP_NTRAIN values are trained. */
#define EQD_MLP_NTRAIN 8833
#define EQD_MLP_NPARAMS 8847

/* Iteration policy of the faster rate-and-state mode (k_fallback > 0): the
 * solve starts from a proposal, stops when the reference's convergence test
 * is met or when the residual is within 16 ulps of the size of its terms
 * (the roundoff floor, where the reference keeps iterating without
 * improving), and after k_fallback iterations without either restarts from
 * the plain guess with the reference's solve (20 iterations).  The proposal
 * is the network's V_prev * exp(r) where it predicts |r| >= gate, else the
 * plain guess; params = NULL (nparams = 0) uses the plain guess throughout.
 * k_fallback = 0 with params = NULL restores the reference's solve.  Takes
 * effect at the next eqd_run. */
int eqd_set_guess(eqd_handle *h, const float *params, int64_t nparams, int32_t k_fallback, float gate);

/* Record (the six network inputs, converged ln V) of every rate-and-state
 * node every `every` steps of the following eqd_run calls into a device
 * buffer of `capacity` rows (7 floats each); capacity = 0 stops and frees it. */
int eqd_harvest(eqd_handle *h, int32_t every, int64_t capacity);
/* Copy up to max_rows recorded rows (row-major [rows][7]); *rows = recorded. */
int eqd_harvest_get(eqd_handle *h, float *out, int64_t max_rows, int64_t *rows);

/* Newton counters accumulated since eqd_commit: stats[0] node solves,
 * stats[1] iterations, stats[2] fallbacks to the plain guess, stats[3] solves
 * that ended without meeting the convergence test, stats[4] solves stopped
 * at the roundoff floor. */
int eqd_newton_stats(eqd_handle *h, int64_t stats[5]);

/* Standalone use of the network on the GPU (host arrays, device `device`):
 * X[n][6] raw inputs, r[n] outputs. */
int eqd_mlp_forward(const float *params, const float *X, int64_t n, float *r_out, int32_t device);
/* eqd_mlp_forward for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  params[EQD_MLP_NPARAMS],
 * X[n][6] and r_out[n] are DEVICE arrays of the caller; the work is enqueued
 * on `stream` (a hipStream_t; null: the null stream) and the call returns
 * without waiting.  It allocates, frees and copies nothing: the network's
 * prepared operands go to a library-owned scratch array of
 * EQD_MLP_NOALLOC_SCRATCH_BYTES on each device it runs on, held until the
 * process ends and shared by every call on that device, so calls that
 * overlap on one device race (serialize them, e.g. on one stream).  n may be
 * any size up to 64 * (2^32 - 1) rows (EQD_EINVAL beyond). */
#define EQD_MLP_NOALLOC_SCRATCH_BYTES 35072
int eqd_mlp_forward_noalloc(const float *params, const float *X, int64_t n, float *r_out, int32_t device,
                            void *stream);
/* Mean squared error of the normalised output against the normalised
 * targets r[n], and its gradient with respect to the first EQD_MLP_NTRAIN
 * parameters, over all n samples. */
int eqd_mlp_grad(const float *params, const float *X, const float *r, int64_t n, float *grad_out,
                 float *loss_out, int32_t device);

typedef struct {
    int32_t epochs, batch;
    float lr, beta1, beta2, eps;
    uint64_t seed; /* shuffling */
    int32_t device;
} eqd_mlp_train_opts;

/* Adam on the mean squared error, batches drawn from a seeded permutation
 * each epoch.  params holds the initial parameters (with the normalisation)
 * and receives the trained ones; epoch_loss[epochs] the mean loss per epoch. */
int eqd_mlp_train(const float *X, const float *r, int64_t n, const eqd_mlp_train_opts *opts, float *params,
                  float *epoch_loss);

/* ---------------------------------------------------------------------------
 * Ground-motion and source-evolution records: the data output_gm and
 * output_src_evol write every 10 steps (src/driver.f90:30-33,
 * src/library_output.f90:267-312) recorded on the device, so the host no
 * longer copies all of velArr and fric back to write a few nodes.  At every
 * step nt with (nt - 1) % every == 0 the velocities of node_ids (1-based, e.g.
 * surfaceNodeIdArr) and fric(47) of every pair go to the next of `capacity`
 * records; capacity 0 stops recording.  The records live in a device ring of
 * `capacity` records; at the end of every run they are copied, on a stream of
 * their own and without stopping the steps, into a pinned host ring of the
 * same size, from which eqd_record_get hands them over: a run writes device
 * slots whose records are already on the host while the host drains the
 * records of the runs before it.  eqd_record_get copies the records since the
 * last call (vel [count][n_nodes][3], slip_rate [count][nf], either may be
 * NULL), in step order, and empties the buffer; with a run in flight
 * (eqd_run_async) the records of that run stay for the next call.  EQD_EINVAL
 * if more records wait than max_records (nothing is copied), or if more
 * record steps passed than the capacity holds (the first `capacity` are
 * copied, the rest were dropped). */
int eqd_record(eqd_handle *h, const int32_t *node_ids, int32_t n_nodes, int32_t every, int32_t capacity);
int eqd_record_get(eqd_handle *h, double *vel, double *slip_rate, int32_t max_records, int32_t *count);

/* The hypocentre dynamics that showSourceDynamics (src/faulting.f90:343-365)
 * prints every step for each pair whose slave node lies within 1e-5 m of
 * (xsource, zsource) in x and z: the library keeps, per step and pair,
 * EQD_HYP_VALS values taken where the reference prints them (after the
 * friction solve): timeElapsed, fric 78, 79, 80 (normal, strike, dip
 * traction), 23, 73, 71, 72 (slips), 74, 75, 76 (slip rates), 77, 1, 2, 3,
 * 9, 10, 20, 81.  eqd_hypocentre_pairs gives the pairs (1-based, pair
 * order, up to cap of them; *n = how many there are).  eqd_hypocentre_get
 * copies steps first_step .. first_step + count - 1 into out
 * [count][n][EQD_HYP_VALS]; the steps must be done (a run in flight is not
 * waited for: its steps are not yet asked for). */
#define EQD_HYP_VALS 19
int eqd_hypocentre_pairs(eqd_handle *h, int32_t *pairs, int32_t cap, int32_t *n);
int eqd_hypocentre_get(eqd_handle *h, int32_t first_step, int32_t count, double *out);

/* ---------------------------------------------------------------------------
 * Device-side one-time precompute: replaces the element loop of
 * assembleGlobalMass (src/assembleGlobalMass.f90:283-406, with
 * calcGlobalShapeFunc, src/calcGlobalShapeFunc.f90:1-75, contm,
 * calcSSPhi4Hrgls and vlm, src/library.f90:60-93): per element dN/dx (eleshp),
 * Jacobian (eledet), lumped mass (elemass), SS and Phi, and each element's
 * lumped mass added to its nodes' equations (nodalMassArr) and to fnms, in
 * the reference's element order (bit-identical sums).  Host arrays in the
 * reference's layouts; nodalMassArr and fnms are added to, as the reference
 * adds to its zeroed arrays.  Returns EQD_EINVAL with *bad_elem (1-based) for
 * a non-positive Jacobian or a node id out of range; *device_ms (may be NULL)
 * receives the kernels' time without the copies. */
typedef struct {
    int64_t nn, ne, neq, size_eqidx;
    const double *meshCoor;            /* (3, nn) */
    const int32_t *nodeElemIdRelation; /* (8, ne), 1-based */
    const int32_t *elemTypeArr;        /* (ne): 11/12 are wedges */
    const double *mat;                 /* (ne, 5): vp, vs, rho, lambda, mu */
    const int32_t *numOfDofPerNodeArr, *eqNumStartIndexLoc, *eqNumIndexArr;
    double *eleshp, *eledet, *elemass, *ss, *phi; /* out: (3,8,ne) (ne) (24,ne) (6,ne) (8,4,ne) */
    double *nodalMassArr, *fnms;                  /* in/out: (neq) (nn) */
    int32_t device;
} eqd_precompute_args;
int eqd_precompute(const eqd_precompute_args *a, int32_t *bad_elem, double *device_ms);

/* ---------------------------------------------------------------------------
 * Launch shapes, captured steps and ensembles.  No counterpart in the
 * reference (one MPI rank per subdomain, one scenario per run); they change
 * how the step is launched, never what it computes: every launch shape, the
 * captured graph and the ensemble give bit-identical results.
 *
 * Launch kinds: */
#define EQD_KIND_ELEMENT 0     /* regular elements */
#define EQD_KIND_ELEMENT_PML 1 /* PML elements */
#define EQD_KIND_NODE 2        /* 3-component nodes */
#define EQD_KIND_NODE_PML 3    /* PML nodes */
#define EQD_KIND_FAULT 4       /* split-node pairs */

/* Block size (64, 128 or 256; fault pairs 64 or 128) and XCD remapping (1:
 * each of the 8 XCDs takes one contiguous range of the launch) of one kind. */
int eqd_set_launch(eqd_handle *h, int32_t kind, int32_t block, int32_t remap);
int eqd_get_launch(eqd_handle *h, int32_t kind, int32_t *block, int32_t *remap);

/* Autotuner: times every launch shape of each kind on the handle's own
 * scenario (1 + steps steps per candidate, state saved before and restored
 * after, so the scenario does not advance; needs nt + 1 + steps <= nstep),
 * keeps the fastest, and writes a JSON report (per kind: the choice, its
 * milliseconds per launch and all candidates) into report[report_cap]. */
int eqd_tune(eqd_handle *h, int32_t steps, char *report, int64_t report_cap);

/* on = 1: eqd_run captures one step as a HIP graph (once) and replays it;
 * on = 0 (the default): eqd_run launches every step's kernels, which measured
 * faster on this ROCm for one scenario (432,000 hexahedra: 0.094 against
 * 0.105 ms per FP32 step; 3.46 M hexahedra: 0.491 against 0.502 ms).
 * Per-kernel timing (eqd_set_timing) runs steps uncaptured.
 * Replayed and launched steps compute the same bits. */
int eqd_set_graph(eqd_handle *h, int32_t on);

/* An ensemble advances committed scenarios of one precision on one device
 * together: each step is one launch per kind over all of them, captured as a
 * HIP graph and replayed (eqd_ensemble_set_graph as eqd_set_graph, default
 * 1).  Members keep their own state (eqd_get, eqd_clock,
 * eqd_newton_stats work as usual) but cannot be run or destroyed on their own
 * until the ensemble is destroyed. */
typedef struct eqd_ensemble eqd_ensemble;
int eqd_ensemble_create(eqd_handle *const *members, int32_t n, eqd_ensemble **out);
/* nsteps steps of every member (each needs nt + nsteps <= its nstep). */
int eqd_ensemble_run(eqd_ensemble *e, int32_t nsteps);
int eqd_ensemble_tune(eqd_ensemble *e, int32_t steps, char *report, int64_t report_cap);
int eqd_ensemble_set_launch(eqd_ensemble *e, int32_t kind, int32_t block, int32_t remap);
int eqd_ensemble_set_graph(eqd_ensemble *e, int32_t on);
int eqd_ensemble_set_timing(eqd_ensemble *e, int32_t on);
int eqd_ensemble_kernel_time(eqd_ensemble *e, const char *name, double *avg_ms, int64_t *launches);
const char *eqd_ensemble_last_error(eqd_ensemble *e);
int eqd_ensemble_destroy(eqd_ensemble *e);

#ifdef __cplusplus
}
#endif
#endif
