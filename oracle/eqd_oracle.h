/* Test infrastructure, not product code.
 *
 * CPU FP64 restatement of EQdyna 5.3.3's explicit time step (the body of
 * src/driver.f90:9-34 and everything it calls), operating on the reference's
 * own data layout (Fortran column-major arrays, 1-based ids, as listed in
 * src/globalvar.f90:78-101).  It is the checker for the HIP path: tests and the
 * CPU-baseline leg of bench.py call it; the product never does.
 */
#ifndef EQD_ORACLE_H
#define EQD_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    double dt, rdampk, rdampm, w, grav, roumax, rhow, gamar;
    double ccosphi, sinphi, tv, slipRateThres;
    double xsource, ysource, zsource, nucR, nucRuptVel, nucdtau0, nucT;
    double kapa_hg, dx, R, vmaxPML, PMLb[8], pi;
    int32_t nstep, nn, ne, neq, nf, non, noff, nPML;
    int32_t C_elastic, C_Q, C_hg, C_nuclea, friclaw, TPV, insertFaultType, nucfault;
    double fric_tp_h;                /* thermal pressurization half-width (module variable) */
} eqdo_scalars;

typedef struct {
    /* mesh (read only) */
    const double *meshCoor;          /* (3, nn) */
    const int32_t *eqNumIndexArr;    /* (sizeOfEqNumIndexArr) */
    const int32_t *eqNumStartIndexLoc; /* (nn) */
    const int32_t *numOfDofPerNodeArr; /* (nn) */
    const int32_t *ien;              /* nodeElemIdRelation (8, ne) */
    const int32_t *elemTypeArr;      /* (ne) */
    const double *mat;               /* (ne, 5) */
    const double *eleshp;            /* (3, 8, ne) */
    const double *eledet;            /* (ne) */
    const double *elemass;           /* (24, ne) */
    const double *ss;                /* (6, ne) */
    const double *phi;               /* (8, 4, ne) */
    const int32_t *stressCompIndexArr; /* (ne) */
    const double *eleporep;          /* (ne) */
    const double *nodalMassArr;      /* (neq) */
    const double *fnms;              /* (nn) */
    const int32_t *nsmp;             /* (2, nf) */
    const double *un, *us, *ud;      /* (3, nf) */
    const double *arn;               /* (nf) */
    const int32_t *anonfs;           /* (3, non) or NULL */
    const int32_t *idhist;           /* (3, 6*noff) or NULL */
    /* state (read/write) */
    double *stressArr, *pstrain, *v1, *velArr, *dispArr, *nodalForceArr, *fric, *fnft;
    /* outputs */
    double *onFault;                 /* (12, nstep, max(non,1)) */
    double *offFault;                /* (6*noff+1, nstep) or NULL */
    /* clock: time step counter and accumulated time, carried across calls */
    int32_t nt;
    double timeElapsed;
    /* error report */
    int32_t err_node, err_nt;
    /* thermal pressurization history onFaultTPHist (2, nf, nstep): slip rate
       and shear traction of every pair at every step (friclaw 5), or NULL */
    double *tpHist;
} eqdo_state;

/* Advance nsteps time steps.  Returns 0, or 1 when a NaN velocity appears
 * (the reference stops at src/driver.f90:147-152; err_node/err_nt say where). */
int eqdo_run(const eqdo_scalars *s, eqdo_state *st, int32_t nsteps);
/* 0: the reference's force grouping (default); 1: the GPU kernels' (see eqd_oracle.c) */
void eqdo_set_grouping(int32_t g);

/* Individual pieces, exposed so tests can pin them in isolation. */
void eqdo_comdampv(const eqdo_scalars *s, double x, double y, double z, double dv[9]);
void eqdo_qconstant(double Q, int32_t k, double pi, double *taok, double *wk, double *c1);
void eqdo_slip_law(double V2, double *psi, const double *fricsgl, double dt, double *xmu, double *dxmudv);
void eqdo_ageing_law(double V2, double *theta, const double *fricsgl, double dt, double *xmu, double *dxmudv);

#ifdef __cplusplus
}
#endif
#endif
