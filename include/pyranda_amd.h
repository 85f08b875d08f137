/*
 * pyranda_amd C-ABI: MI355X-native compact finite-difference core.
 *
 * Drop-in boundary for pyranda's f2py module `parcop` (pyranda/parcop/parcop.f90,
 * reached only through pyranda/pyrandaMPI.py:151-155,664-740 and
 * pyranda/pyrandaMesh.py:85-208). Every entry point below names the parcop
 * routine it replaces.
 *
 * Conventions
 *  - Fields are DEVICE pointers to (nx,ny,nz) arrays in the reference's
 *    Fortran order: x fastest, then y, then z (parcop's `val(nx,ny,nz)`).
 *  - Outputs are caller-allocated (parcop returns fresh intent(out) arrays;
 *    the Python shim allocates them). Inputs and outputs must not alias.
 *  - Work is enqueued on the context's stream (pyr_set_stream); calls do not
 *    synchronise the host.
 *  - Errors return a nonzero status; pyr_last_error() gives the message
 *    (replaces the reference's print + STOP, compact_d1.f90:65-72).
 *  - `_f64` entry points compute like the reference in FP64. `_r32` entry
 *    points read/write FP32 fields, sweep in FP32 and refine once with an
 *    FP64 residual against the same banded operator.
 */
#ifndef PYRANDA_AMD_H
#define PYRANDA_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pyr_ctx pyr_ctx;

/* boundary codes per side, as the strings pyrandaMPI passes to parcop.setup
 * (pyrandaMPI.py:101-131): 'PERI', 'NONE', 'SYMM' */
enum { PYR_BC_PERI = 0, PYR_BC_NONE = 1, PYR_BC_SYMM = 2 };

/* operator ids for the generic entry point */
enum { PYR_OP_D1 = 0, PYR_OP_D2 = 1, PYR_OP_D8 = 2, PYR_OP_SFILTER = 3, PYR_OP_GFILTER = 4 };
/* or-ed into an operator id: the antisymmetric member of the SYMM pair
 * (compact.f90:78-84) where the direction has a SYMM end, as
 * d1x/d8x(v, dv, bc=-1) select it (compact_operators.f90:300-303); elsewhere
 * the default operator */
enum { PYR_OP_ANTISYM = 0x10 };

/* how a directional result lands in `out` (fuses the pointwise combination of
 * lap/div/ring into the operator's store) */
enum { PYR_EPI_STORE = 0, PYR_EPI_ADD = 1, PYR_EPI_ABS_SCALED = 2, PYR_EPI_MAXABS_SCALED = 3 };

/* precision of the generic entry point */
enum { PYR_F64 = 0, PYR_R32 = 1 };
/* mixed precisions of the RK update (F and U of different widths): FP32
 * right-hand side with an FP64 field, FP64 right-hand side with an FP32
 * field (PYR_F64: both double, PYR_R32: both float) */
enum { PYR_RK_F32_U64 = 2, PYR_RK_F64_U32 = 3 };

/* --- setup / teardown -------------------------------------------------- */

/* Replaces parcop.setup (parcop.f90:23-61) for one Cartesian block on one GPU.
 * x1..zn are node bounds as pyranda gives them; bc6 = {x1,xn,y1,yn,z1,zn}. */
int pyr_setup(pyr_ctx** ctx, int nx, int ny, int nz, double x1, double xn, double y1,
              double yn, double z1, double zn, const int* bc6);
/* Same tables built on the host only (no device allocation): for CPU tests of
 * the host-side plan and pyr_host_emulate. */
int pyr_setup_host(pyr_ctx** ctx, int nx, int ny, int nz, double x1, double xn, double y1,
                   double yn, double z1, double zn, const int* bc6);
int pyr_destroy(pyr_ctx* ctx);
/* HIP stream (hipStream_t as void*), 0 = null stream */
int pyr_set_stream(pyr_ctx* ctx, void* stream);

/* Solver of periodic implicit lines. No counterpart in the reference, which
 * always uses the bordered cyclic LU (pentadiagonal.f90:64-131, 654-825):
 * PYR_PSOLVE_CIRCULANT (default) factors the circulant LHS as c Q^T Q and
 * runs constant-coefficient recurrences; PYR_PSOLVE_LU uses the chunked LU
 * of bounded lines with the cyclic coupling in its reduced system. Both give
 * the same solution to rounding. */
enum { PYR_PSOLVE_CIRCULANT = 0, PYR_PSOLVE_LU = 1 };
int pyr_set_periodic_solver(pyr_ctx* ctx, int mode);

/* Arithmetic of the solves on FP32 fields (PYR_R32 entry points and
 * pyr_apply with PYR_R32). No counterpart in the reference (FP64 only):
 * PYR_R32_SOLVE_REFINED (default) solves in FP32 and applies one FP64-residual
 * refinement; PYR_R32_SOLVE_FP64 forms the RHS and solves in FP64 registers
 * and rounds once to FP32 (on gfx950 FP32 and FP64 vector FMAs issue at the
 * same rate, so this is the cheaper of the two). */
enum { PYR_R32_SOLVE_REFINED = 0, PYR_R32_SOLVE_FP64 = 1 };
int pyr_set_r32_solve(pyr_ctx* ctx, int mode);
/* The same choice for one operator (PYR_OP_*) along one direction, over the
 * context's (-1 restores it); composites (filters, ring, Laplacian, grad,
 * divergence) follow their operator per direction. Set by pyranda_amd's
 * autotuner (ParcopContext.autotune with solve_modes). No counterpart in the
 * reference. */
int pyr_set_solve_tuning(pyr_ctx* ctx, int op, int dir, int r32_solve);
const char* pyr_last_error(void);

/* --- mesh (parcop.f90 xgrid/ygrid/zgrid/dxgrid/mesh_getgridlen; mesh.f90:173-212)
 * Host outputs: coords of one axis (length n of that axis), the patch spacing
 * d[3] (metric of ddx), the mesh lengths d1..d3 (ring scale) and GridLen. */
int pyr_axis_coords(pyr_ctx* ctx, int axis, double* coords_host);
int pyr_mesh_scalars(pyr_ctx* ctx, double* dpatch3, double* dmesh3, double* gridlen);

/* --- FP64 operators (device pointers) ----------------------------------
 * divergence and pRingV differentiate the normal component of each
 * direction with the antisymmetric member of the SYMM pair where that
 * direction has a SYMM end (ddx(fx,..,isymX), ringx(f,..,isymX);
 * operators.f90:38-53,661-671); every other operator uses the default. */
int pyr_ddx_f64(pyr_ctx* ctx, const double* val, double* dval);      /* parcop.f90:225 ddx */
int pyr_ddy_f64(pyr_ctx* ctx, const double* val, double* dval);      /* parcop.f90:235 ddy */
int pyr_ddz_f64(pyr_ctx* ctx, const double* val, double* dval);      /* parcop.f90:245 ddz */
int pyr_dd8x_f64(pyr_ctx* ctx, const double* val, double* dval);     /* parcop.f90:283 dd8x */
int pyr_dd8y_f64(pyr_ctx* ctx, const double* val, double* dval);     /* parcop.f90:290 dd8y */
int pyr_dd8z_f64(pyr_ctx* ctx, const double* val, double* dval);     /* parcop.f90:297 dd8z */
int pyr_d2x_f64(pyr_ctx* ctx, const double* val, double* dval);      /* compact_operators.f90:130 d2x */
int pyr_d2y_f64(pyr_ctx* ctx, const double* val, double* dval);      /* compact_operators.f90:156 d2y */
int pyr_d2z_f64(pyr_ctx* ctx, const double* val, double* dval);      /* compact_operators.f90:182 d2z */
int pyr_plaplacian_f64(pyr_ctx* ctx, const double* val, double* dval); /* parcop.f90:303 plaplacian */
int pyr_pring_f64(pyr_ctx* ctx, const double* val, double* dval);    /* parcop.f90:313 pRing (L=2) */
int pyr_pringv_f64(pyr_ctx* ctx, const double* vx, const double* vy, const double* vz,
                   double* dval);                                    /* parcop.f90:324 pRingV (L=1) */
int pyr_sfilter_f64(pyr_ctx* ctx, const double* val, double* dval);  /* parcop.f90:337 sFilter (fbar) */
int pyr_gfilter_f64(pyr_ctx* ctx, const double* val, double* dval);  /* parcop.f90:348 gFilter (gbar) */
int pyr_gfilterdir_f64(pyr_ctx* ctx, const double* val, double* dval, int dir); /* parcop.f90:359 */
int pyr_grads_f64(pyr_ctx* ctx, const double* val, double* v1, double* v2,
                  double* v3);                                       /* parcop.f90:371 gradS */
int pyr_divergence_f64(pyr_ctx* ctx, const double* fx, const double* fy, const double* fz,
                       double* val);                                 /* parcop.f90:202 divergence */

/* --- FP32 fields, FP32 sweeps + one FP64-residual refinement ------------ */
int pyr_ddx_r32(pyr_ctx* ctx, const float* val, float* dval);
int pyr_ddy_r32(pyr_ctx* ctx, const float* val, float* dval);
int pyr_ddz_r32(pyr_ctx* ctx, const float* val, float* dval);
int pyr_plaplacian_r32(pyr_ctx* ctx, const float* val, float* dval);
int pyr_pring_r32(pyr_ctx* ctx, const float* val, float* dval);
int pyr_sfilter_r32(pyr_ctx* ctx, const float* val, float* dval);
int pyr_gfilter_r32(pyr_ctx* ctx, const float* val, float* dval);
int pyr_gfilterdir_r32(pyr_ctx* ctx, const float* val, float* dval, int dir);
int pyr_grads_r32(pyr_ctx* ctx, const float* val, float* v1, float* v2, float* v3);
int pyr_divergence_r32(pyr_ctx* ctx, const float* fx, const float* fy, const float* fz,
                       float* val);

/* --- generic directional operator (compact_operators.f90:13-497) --------
 * op: PYR_OP_*, dir: 0..2, prec: PYR_F64 (double*) or PYR_R32 (float*),
 * epi: PYR_EPI_*, epi_scale multiplies |y| for the ABS epilogues. */
int pyr_apply(pyr_ctx* ctx, int op, int dir, int prec, const void* in, void* out, int epi,
              double epi_scale);

/* --- fbar that allocates nothing -----------------------------------------
 * No counterpart in the reference. For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc. The filter of pyr_sfilter_f64 (prec PYR_F64) or
 * pyr_sfilter_r32 (prec PYR_R32), results identical, with the field
 * temporary of its x -> y -> z sweeps in a file-scope device array of 64 MiB
 * (one per GPU, held until the process ends; its address is taken at
 * pyr_setup): fields of at most 64 MiB (8,388,608 FP64 or 16,777,216 FP32
 * points, e.g. 256^3 in FP32); larger fields return an error and launch
 * nothing. Every call on a GPU uses that GPU's one array: calls that overlap
 * on one GPU (other streams, other threads) race. */
int pyr_sfilter_noalloc(pyr_ctx* ctx, int prec, const void* val, void* dval);

/* --- FP64 operators that store the FP32 rounding of their result --------
 * For FP32 runs that keep some fields in FP64: the conservative filter of an
 * FP64 state (the field is its rounding), and operators of FP64 fields whose
 * result is stored in FP32 (e.g. a velocity derivative). No counterpart in
 * the reference (FP64 only); the operators are those of
 *   pyr_sfilter_f64_rounded: parcop.f90:337 sFilter (fbar), out32[i] = (float)dval[i]
 *   pyr_apply_rounded:       compact_operators.f90:13-497, one sweep of an
 *                            FP64 field with PYR_EPI_STORE, FP64 result in out
 *                            and its rounding in out32
 *   pyr_apply_to_r32:        the same sweep storing only the rounding (no FP64
 *                            result is written: 8 + 4 bytes per point)
 * out32 holds as many floats as the field has points. */
int pyr_sfilter_f64_rounded(pyr_ctx* ctx, const double* val, double* dval, float* out32);
int pyr_apply_rounded(pyr_ctx* ctx, int op, int dir, const double* in, double* out, float* out32);
int pyr_apply_to_r32(pyr_ctx* ctx, int op, int dir, const double* in, float* out32);

/* --- one RK4 stage update of a conserved variable (pyranda.py:800-804) ----
 * PHI <- dt*F + A*PHI (in place, FP64; with A == 0 PHI is only written), then
 *   PYR_F64: U_out <- U_in + B*PHI                (double fields)
 *   PYR_R32: U64 <- U64 + B*PHI (FP64 state), U_out <- (float)U64; with
 *            U64 == NULL: U_out <- (float)((double)U_in + B*PHI)
 *   PYR_RK_F32_U64 / PYR_RK_F64_U32: as above with F float and U double /
 *            F double and U float (an FP32 run with some fields kept in FP64)
 * F, U_in, U_out have the precision prec names; n = number of points. With U64
 * given, U_out may be NULL: only the FP64 state advances (for a caller whose
 * next read of U is the FP64 state, e.g. the filter of the state). */
int pyr_rk_update(pyr_ctx* ctx, int prec, long long n, const void* F, double* PHI,
                  const void* U_in, void* U_out, double* U64, double dt, double A, double B);
/* The same update with dt read from device memory (one FP64 value) when the
 * kernel runs, so a time step whose dt comes from device reductions needs no
 * host round trip and can be captured in a HIP graph and replayed. No
 * counterpart in the reference (its dt is a host float, pyranda.py:758). */
int pyr_rk_update_dt(pyr_ctx* ctx, int prec, long long n, const void* F, double* PHI,
                     const void* U_in, void* U_out, double* U64, const double* dt, double A,
                     double B);

/* Tuning of one operator along one direction at one precision (for FP32
 * fields: in the context's current FP32 solve mode): x_lds_kb is the LDS
 * budget of an x-line workgroup (8..160 KB; it sets the lines per
 * workgroup), yz_pipe switches the pipelined kernels (0/1; on x also 2: the
 * next group of lines prefetched into registers, used by the explicit
 * filter; on y/z also 3: the plain kernel with runtime row counts, fewer
 * registers), yz_lines the
 * lines per plain y/z workgroup (16, 32 or 64); -1 keeps the context default.
 * Set by pyranda_amd's autotuner (ParcopContext.autotune), which times the
 * candidates on the device. No counterpart in the reference. */
int pyr_set_tuning(pyr_ctx* ctx, int op, int dir, int prec, int x_lds_kb, int yz_pipe,
                   int yz_lines);

/* --- groups: one operator over several simulations' fields ---------------
 * No counterpart in the reference (pyranda advances one simulation per MPI
 * communicator, pyranda/pyranda.py:32-100). The members are contexts of any
 * grid sizes (pyr_setup each) whose fields are stored in one allocation,
 * member m's (nx,ny,nz) block at offsets[m] elements; offsets that are
 * multiples of 16 elements keep the 16-byte row copies. pyr_group_apply is
 * pyr_apply on every member at once: one launch for all members' lines of
 * the direction (any lengths 8..512), plus a null-operator pass for members
 * whose direction has fewer than 4 points. The epilogue constant of member m is
 * epi_scale * d_m^mesh_power (d_m its mesh spacing d1..d3 along the
 * direction: mesh_power 2 gives the ring scaling of operators.f90:615-643).
 * Periodic lines and explicit filters only (the bounded LU is not grouped).
 * The first call of each (op, dir, prec, epi, scale) builds and uploads the
 * members' line arguments (synchronous copy); later calls only launch. */
typedef struct pyr_group pyr_group;
int pyr_group_create(pyr_group** group, int n_members, pyr_ctx* const* members,
                     const long long* offsets);
int pyr_group_destroy(pyr_group* group);
int pyr_group_set_stream(pyr_group* group, void* stream);
int pyr_group_apply(pyr_group* group, int op, int dir, int prec, const void* in, void* out,
                    int epi, double epi_scale, int mesh_power);
/* RK4 stage update of a group's fields (n points in all) with each member's
 * own dt: dt_member[member_of_point[i]] (FP64, device memory). */
int pyr_rk_update_member_dt(void* stream, int prec, long long n, const void* F, double* PHI,
                            const void* U_in, void* U_out, double* U64, const double* dt_member,
                            const int* member_of_point, double A, double B);

/* --- fused elementwise kernels for EOM pointwise algebra ----------------
 * No counterpart in the reference (its pointwise algebra is numpy, pyrandaEq.py:39).
 * Source is generated by pyranda_amd/fuse.py: one kernel taking a by-value
 * argument block {long long n; void* p[PYR_PW_MAX_PTRS]; double s[PYR_PW_MAX_SCALARS];
 * int vec;}; the launch sets vec to 1 when every pointer passed is 16-byte
 * aligned (the generated elementwise loops then move 16-byte vectors). */
#define PYR_PW_MAX_PTRS 24
#define PYR_PW_MAX_SCALARS 16
int pyr_pw_compile(const char* source, const char* kernel_name, void** fn);
int pyr_pw_check_source(const char* source, const char* kernel_name);
/* Workgroups of PYR_PW_BLOCK threads; pyr_pw_launch picks the grid,
 * pyr_pw_launch_blocks takes it (reduction kernels write one partial per
 * workgroup). */
#define PYR_PW_BLOCK 256
int pyr_pw_launch(void* fn, void* stream, long long n, int nptr, void* const* ptrs, int nscal,
                  const double* scal);
int pyr_pw_launch_blocks(void* fn, void* stream, long long n, int blocks, int nptr,
                         void* const* ptrs, int nscal, const double* scal);
const char* pyr_pw_last_error(void);

/* --- learned LAD closure (BF16 MFMA) -------------------------------------
 * No counterpart in the reference: pyranda evaluates the LAD artificial
 * properties with compact operators only (examples/RM3D.py:95-104). A
 * pointwise MLP maps the 3x3 neighbourhoods of three feature fields of a 2-D
 * (nx, ny) grid (x fastest; neighbours wrap on periodic axes and are clamped
 * otherwise) to two outputs: 27 inputs (padded to 32) -> 64 -> 64 -> 2, ReLU
 * after the hidden layers, BF16 weights and activations, FP32 accumulation.
 * Input k = 9 f + 3 (dy + 1) + (dx + 1) of point (i, j) is
 * (feature_f(i+dx, j+dy) - fshift[f]) * fscale[f].
 *
 * `params` is one device block of PYR_MLP_PARAM_BYTES bytes:
 *   bf16 W1[32][64], W2[64][64], W3[64][16]   (row = input unit; rows 27..31
 *                                                of W1 and columns 2..15 of W3 are 0)
 *   at PYR_MLP_PARAM_FLOAT_OFFSET, float b1[64], b2[64], b3[16], fshift[4],
 *                                     fscale[4], oscale[4]
 * All fields of one call (features, targets, outputs) have the precision
 * `prec` and the (nx, ny) layout.
 * Inference: out_j = max(o_j, 0) * oscale[j].
 * Training: the gradient of L = sum_points sum_j (o_j - t_j / oscale[j])^2 / n
 * with respect to every parameter, in PYR_MLP_GRAD_FLOATS floats:
 *   gW1[32][64], gW2[64][64], gW3[64][16], gb1[64], gb2[64], gb3[16], then
 *   sum_points sum_j (o_j - t_j / oscale[j])^2 (not divided by n).
 * `partial` is caller-provided workspace of npartial * PYR_MLP_GRAD_FLOATS
 * floats (one block per workgroup, at most npartial workgroups; the sum over
 * blocks is taken in a fixed order). Both calls enqueue on `stream` (a
 * hipStream_t) and return without synchronising. */
#define PYR_MLP_PARAM_FLOAT_OFFSET 14336
#define PYR_MLP_PARAM_BYTES 14960
#define PYR_MLP_GRAD_FLOATS 7313
int pyr_mlp_infer(void* stream, int prec, int nx, int ny, int periodic_x, int periodic_y,
                  const void* const* features, const void* params, void* const* out);
/* pyr_mlp_infer plus a per-point sensor: with sensor > 0 a point is flagged
 * when the largest |input| of its 27 normalised inputs exceeds `sensor`;
 * mask (one byte per point, when not NULL) receives 1 for a flagged point
 * and 0 otherwise, and the number of flagged points is added to *count (when
 * not NULL; the caller clears it). The outputs are written at every point. */
int pyr_mlp_infer_sensed(void* stream, int prec, int nx, int ny, int periodic_x, int periodic_y,
                         const void* const* features, const void* params, void* const* out, float sensor,
                         unsigned char* mask, int* count);
int pyr_mlp_train_grad(void* stream, int prec, int nx, int ny, int periodic_x, int periodic_y,
                       const void* const* features, const void* const* targets, const void* params,
                       float* partial, int npartial, float* grad);

/* --- host-side checks (no GPU needed) ----------------------------------- */
/* Weight tables as the library restates them (stencils.f90), for comparison
 * with the oracle: ali[5], ari[9], alb1/alb2[3*4*5], arb1/arb2[3*4*9]. */
int pyr_host_weights(int op, double* ali, double* ari, double* alb1, double* alb2, double* arb1,
                     double* arb2, int* dims6);
/* Runs the GPU kernels' per-thread arithmetic (chunked line solve) on the host
 * over every line of one direction: a CPU emulation of the device code path
 * used only by tests to check the partitioned algorithm without a GPU. */
int pyr_host_emulate(pyr_ctx* ctx, int op, int dir, int prec, const void* in, void* out);

#ifdef __cplusplus
}
#endif
#endif
