/*
 * CPU oracle for the compact finite-difference core (TEST INFRASTRUCTURE).
 *
 * This is a plain-C FP64 restatement of the reference algorithm, used only by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
 * checker. The product path (pyranda_amd/) never links or loads it.
 *
 * What it restates (reference file:line, paths relative to the pyranda repo):
 *   weights      pyranda/parcop/stencils.f90:207-277 (c10d1), :358-428 (c10d2),
 *                :515-621 (c10d8), :713-835 (c8ff8), :1387-1476 (cgfs4),
 *                symmetry folding :2390-2453
 *   assembly     pyranda/parcop/compact_basetype.f90:124-148
 *   LU factor    pyranda/parcop/pentadiagonal.f90:18-39 (bpentLUD1),
 *                :64-131 (ppentLUD1)
 *   line solves  pyranda/parcop/pentadiagonal.f90:629-684 (bpentLUS3x/ppentLUS3x)
 *   RHS forms    pyranda/parcop/compact_d1.f90:100-227, compact_r3.f90:73-157,
 *                compact_r4.f90:79-233
 *   metric       pyranda/parcop/compact_operators.f90:41-45 (/dx), :152 (/dx**2)
 *   composites   pyranda/parcop/operators.f90:38-53 (div), :184-212 (grad),
 *                :522-526 (lap), :615-699 (ring/ringV), :755-778 (gfilter dir),
 *                :781-896 (filter x->y->z)
 *
 * Parity pins (see tests/): weights and LU are compared bit-for-bit with the
 * reference's own stencils.f90/pentadiagonal.f90 compiled into oracle/_ref;
 * periodic operators against their closed-form transfer functions; bounded
 * operators against dense numpy solves; the whole chain against the
 * reference-held unit-test value 0.017490746954566275.
 */
#ifndef COMPACT_ORACLE_H
#define COMPACT_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* operator kinds (compact.f90:26-31 defaults: d1spec=1, d2spec=1, d8spec=1,
 * sfspec=2 -> c8ff8, gfspec=6 -> cgfs4) */
enum { ORA_D1 = 0, ORA_D2 = 1, ORA_D8 = 2, ORA_SF = 3, ORA_GF = 4, ORA_NOPS = 5 };

/* boundary codes per side: PERI, NONE (variant 0), SYMM (variant +1 / -1) */
enum { ORA_BC_PERI = 0, ORA_BC_NONE = 1, ORA_BC_SYMM = 2 };

/* Weight table of one scheme, 0-based; variant index v+1 for v in {-1,0,1}. */
typedef struct {
  int nol, nor, ncl, ncr, nbc1, nbc2, null_option, implicit_op;
  double ali[5];
  double ari[9];
  double alb1[3][4][5], alb2[3][4][5]; /* [variant+1][row][col] */
  double arb1[3][4][9], arb2[3][4][9];
} ora_weight;

void ora_get_weight(int op, ora_weight* w);

/* LU factorisations and solves, column-major c(n,5) / c(n,9) as Fortran. */
void ora_bpentLUD1(double* c, int n);
void ora_ppentLUD1(double* c, int n);
void ora_bpentLUS1(const double* c, double* r, int n);
void ora_ppentLUS1(const double* c, double* r, int n);

/* The per-line sweep used by bpentLUS3x/ppentLUS3x (pentadiagonal.f90:629-684). */
void ora_line_solve(int periodic, const double* c, double* r, int n);

/* Assemble the band (al(n,nal) column-major, ar(nr,n) column-major) and
 * factor it exactly as setup_compact_op1 does for np=1. Returns nal. */
int ora_assemble(int op, int n, int periodic, int bc1, int bc2, double* al, double* ar);

typedef struct ora_ctx ora_ctx;

/* nx,ny,nz and spacings; bc codes per side (x1,xn,y1,yn,z1,zn). */
ora_ctx* ora_create(int nx, int ny, int nz, double dx, double dy, double dz,
                    const int* bc6);
void ora_destroy(ora_ctx* c);

/* One directional operator with the reference's metric scaling and +v for
 * filters: dir 0=x 1=y 2=z. v and dv are (nx,ny,nz) column-major. */
int ora_apply(ora_ctx* c, int op, int dir, const double* v, double* dv);
/* isym=-1: antisymmetric member of the SYMM pair where the direction has one
   (compact_operators.f90:300-303); +1: the default operator */
int ora_apply_sym(ora_ctx* c, int op, int dir, int isym, const double* v, double* dv);

/* Composites as parcop.f90 exposes them. */
int ora_sfilter(ora_ctx* c, const double* v, double* out);   /* fbar */
int ora_gfilter(ora_ctx* c, const double* v, double* out);   /* gbar */
int ora_gfilterdir(ora_ctx* c, const double* v, double* out, int dir1); /* 1..3 */
int ora_ring(ora_ctx* c, const double* v, double* out);      /* pRing, L=2 */
int ora_ringv(ora_ctx* c, const double* vx, const double* vy, const double* vz,
              double* out);                                  /* pRingV, L=1 */
int ora_laplacian(ora_ctx* c, const double* v, double* out);
int ora_divergence(ora_ctx* c, const double* fx, const double* fy,
                   const double* fz, double* out);
int ora_grads(ora_ctx* c, const double* v, double* gx, double* gy, double* gz);

/* OpenMP thread count used by ora_apply (1 = serial). */
void ora_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
