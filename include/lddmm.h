/*
 * lddmm.h -- C ABI of the MI355X (gfx950) point-set LDDMM kernel library
 * (liblddmm_hip.so, built from registration_amd/csrc/lddmm_kernels.hip).
 *
 * All arrays are device pointers to row-major float64 data unless stated
 * otherwise: a point set x(n, d) is x[k*d + c].  Every entry point is
 * stream-ordered on the hipStream_t passed as `stream` (NULL = the null
 * stream), returns LDDMM_OK or a negative error code, and never allocates
 * device memory: scratch space comes in through (work, work_bytes), sized by
 * the matching *_workspace_bytes query.
 *
 * Precision (`prec`):
 *   LDDMM_PREC_FP64   pairwise arithmetic in FP64 (the plain FP64-equivalent form);
 *   LDDMM_PREC_FP32C  pairwise arithmetic in FP32 with Kahan-compensated
 *                     accumulation across source tiles; per-chunk partials and
 *                     all totals in FP64.
 *   LDDMM_PREC_FP64E  the data-term energies of FP32C registrations: FP64 pairwise
 *                     arithmetic and accumulation as LDDMM_PREC_FP64, with pairs
 *                     whose Gaussian kernel value is below 2^-40 culled (2^-56 in
 *                     FP64) and the exponential of those below 2^-16 taken in FP32
 *                     (2^-26 in FP64).  Energy-only segments only (KNORM, and
 *                     CURRENTS, MEASURE, VARIFOLD without per-row outputs); other
 *                     segments are rejected.  Sums of a shape against a nearby
 *                     one lose the culled terms from both sides alike, so their
 *                     difference (the data term) keeps far more digits than the
 *                     single sums: on the 81,920-face surface pair it stays within
 *                     2.2e-10 of FP64 at a data term 6.4e-7 of the sums.
 *
 * Kernel selection (`kernel`, `order`) follows kernelFunctions.KernelSpec
 * (py-lddmm/kernelFunctions.py:238-280): LDDMM_KERNEL_GAUSS is
 * K(r) = exp(-r^2 / (2 sigma^2)); LDDMM_KERNEL_LAPLACIAN of order 0..4 is
 * P_order(r/sigma) exp(-r/sigma) with the float32-rounded coefficient tables of
 * py-lddmm/kernelFunctions.f90:25-26,85-86,461.
 */
#ifndef LDDMM_H
#define LDDMM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  LDDMM_OK = 0,
  LDDMM_EINVAL = -1,     /* bad size, dimension, kernel or order */
  LDDMM_EWORKSPACE = -2, /* work_bytes smaller than the query returned */
  LDDMM_EHIP = -3,       /* a HIP runtime call failed (see lddmm_last_hip_error) */
  LDDMM_EUNSUPPORTED = -4,
  LDDMM_ESCRATCH = -5    /* request larger than lddmm_apply_k_noalloc's fixed scratch */
};

enum { LDDMM_KERNEL_GAUSS = 0, LDDMM_KERNEL_LAPLACIAN = 1 };
enum { LDDMM_PREC_FP64 = 0, LDDMM_PREC_FP32C = 1, LDDMM_PREC_FP64E = 2 };

/* Operation codes of the segmented pair-sum engine (lddmm_segments_run). */
enum {
  /* out0[k,:] = sum_l K(y_k, x_l) w0[l,:]                        (applyK) */
  LDDMM_OP_APPLY_K = 0,
  /* out0[k,:] = sum_l 2 K'(r^2) (y_k - x_l) sum_j <tw_j[k], sw_j[l]>
   * over j < na                                             (applyDiffKT) */
  LDDMM_OP_DIFF_KT = 1,
  /* one explicit Euler step of the landmark flow, targets = sources = y:
   *   v = K(y, y) w0;  out0 = y + dt v [+ dt (A y + b)];  out1 = v (if non-NULL);
   *   rowdot[k] = <w0[k], v[k]>;  total[0] = sum_k rowdot[k]
   * with the optional affine part A = w1 (d x d row-major), b = w2 (d), both or
   * neither (pointEvolution.py:697-699)                       */
  LDDMM_OP_FLOW_STEP = 2,
  /* one backward step of the Hamiltonian covector, targets = sources = y,
   * a = w0, p = w1:
   *   out0 = p + dt sum_l 2K'(r^2)(y_k - y_l)[<p_k - 2 lambda a_k, a_l> + <a_k, p_l>]
   *          [+ dt p A]
   * with lambda = coef and the optional affine matrix A = w2 (d x d row-major,
   * pointEvolution.py:762-763)                                  */
  LDDMM_OP_COVECTOR_STEP = 3,
  /* RKHS products: u = K(y, y) w0; total[j] = sum_k <g_j[k], u[k]>, j < ndot */
  LDDMM_OP_KDOT = 4,
  /* currents data term of a deformed shape (c1, cr1) = (y, w0) against a
   * target (c2, cr2) = (x1, w1):
   *   total[0] = <cr1, K(c1,c1) cr1> - 2 <cr1, K(c1,c2) cr2>   (currentNormDef)
   * and, when out0/out1 are non-NULL (gradient form):
   *   out0 = z1 = K(c1,c1) cr1 - K(c1,c2) cr2
   *   out1 = coef * [applyDiffKT(c1,[cr1],[cr1]) - applyDiffKT(c2,[cr1],[cr2],c1)] */
  LDDMM_OP_CURRENTS = 5,
  /* out0[k,:] = sum_l 2<w0_k, w1_l>(2K''<dx,dp>dx + K' dp), p = w2, dx = y_k - y_l,
   * dp = p_k - p_l (na = 0: applyDDiffK11and12, kernelFunctions.f90:430-496);
   * dp = p_k (na = 1: applyDDiffK11) or -p_l (na = 2: applyDDiffK12), the dense
   * siblings kernelFunctions.py:459-506, with their exact-double Laplacian
   * polynomials (:75-104) */
  LDDMM_OP_DDIFF_K11_12 = 6,
  /* total[0] = <w0, K(y, y) w0>  (currentNorm0 when w0 = surfel) */
  LDDMM_OP_KNORM = 7,
  /* measure data term of a deformed shape with centers c1 = y, face weights
   * a1 = w0 (ny) and normals n1 = w2 (ny, d), against a target with centers
   * c2 = x1 and weights a2 = w1 (nx1; nx1 = 0 leaves the target out):
   *   total[0] = <a1, K(c1,c1) a1> - 2 <a1, K(c1,c2) a2>     (measureNormDef;
   *              measureNorm0 of a shape when nx1 = 0)
   * and, when out0/out1 are non-NULL (gradient form, w2 required):
   *   out0 = z1 = (K(c1,c1) a1 - K(c1,c2) a2)[k] * n1[k] / a1[k]
   *   out1 = coef * [applyDiffKT(c1,[a1],[a1]) - applyDiffKT(c2,[a1],[a2],c1)]
   * (surfaces.py:1092-1182, curves.py:621-701; a = sqrt(|normal|^2 + 1e-10) is
   * computed by the caller, e.g. lddmm_geometry_batch's weights output). */
  LDDMM_OP_MEASURE = 8,
  /* varifold data term of a deformed surface (d = 3 only): centers c1 = y,
   * weights a1 = w0 (ny), normals n1 = w2 (ny, 3), against a target with
   * centers c2 = x1, weights a2 = w1 (nx1) and normals n2 = sw[0] (nx1, 3)
   * (nx1 = 0 leaves the target out).  With u = n / a and g = <u_k, u_l>:
   *   total[0] = sum_kl K(c1_k,c1_l) a1_k a1_l (1 + g^2)
   *            - 2 sum_kl K(c1_k,c2_l) a1_k a2_l (1 + g^2)       (varifoldNormDef;
   *              varifoldNorm0 of a shape when nx1 = 0)
   * and, when out1 (and optionally out0) are non-NULL (gradient form):
   *   out0 = z1 = sum_l K a_l (2 g u_l + (1 - g^2) u1_k)
   *   out1 = coef * sum_l 2K'(r^2)(c1_k - x_l) a1_k a_l (1 + g^2)
   * where l runs over the shape's own faces and, with sign -1, the target's.
   * (surfaces.py:1184-1276 with kernelFunctions.f90:499-610; the (F, F) pair
   * weights are formed per pair, never stored). */
  LDDMM_OP_VARIFOLD = 9,
  /* divergence of v = K(., x0) w0 at the targets (x0 NULL means y):
   *   out0[k] = sum_l 2K'(r^2) <w0_l, y_k - x0_l>                (applyDivergence,
   * kernelFunctions.py:510-530; Laplacian orders 1..4 with the dense path's
   * exact-double coefficients, kernelFunctions.py:75-88).  w0: (nx0, d), out0: (ny). */
  LDDMM_OP_DIVERGENCE = 10,
  /* applyDiffK family (kernelFunctions.f90:116-290), targets = sources = y,
   * a1 = w0, a2 = w1 (n, d):  out0[k,:] = sum_l 2K'(r^2) <y_k - y_l, b_kl> a2_l with
   * b_kl = a1_k (na = 1: applyDiffK), -a1_l (na = 2: applyDiffK2) or a1_k - a1_l
   * (na = 3: applyDiffK1and2) */
  LDDMM_OP_DIFF_K = 11
};

/*
 * One segment of a pair-sum launch: one independent kernel sum (one
 * registration, one time step, or one call).  The caller fills the fields
 * marked [in]; lddmm_segments_plan fills the fields marked [plan].
 * Pointers that an operation does not use may be NULL.
 */
typedef struct lddmm_seg {
  const double *y;     /* [in] targets (ny, d)                                  */
  const double *x0;    /* [in] sources of set 0 (nx0, d); NULL means y          */
  const double *x1;    /* [in] sources of set 1 (nx1, d) (currents target)      */
  const double *w0;    /* [in] weights, meaning per operation (see op codes)    */
  const double *w1;    /* [in]                                                   */
  const double *w2;    /* [in]                                                   */
  const double *tw[4]; /* [in] DIFF_KT target-side weights a1[j] (ny, m)        */
  const double *sw[4]; /* [in] DIFF_KT source-side weights a2[j] (nx0, m)       */
  const double *g[2];  /* [in] KDOT second operands (ny, m)                     */
  double *out0;        /* [in] outputs, meaning per operation                   */
  double *out1;        /* [in]                                                   */
  double *rowdot;      /* [in] optional per-row FP64 products (ny)              */
  double *total;       /* [in] FP64 scalar outputs (1..2 per segment)          */
  int64_t ny, nx0, nx1;/* [in] row counts                                        */
  double sigma;        /* [in] kernel width                                      */
  double dt;           /* [in] time step (FLOW_STEP, COVECTOR_STEP)              */
  double coef;         /* [in] lambda (COVECTOR_STEP) or gradient factor (CURRENTS) */
  int32_t kernel;      /* [in] LDDMM_KERNEL_*                                    */
  int32_t order;       /* [in] Laplacian order 0..4 (ignored for the Gaussian)   */
  int32_t m;           /* [in] weight columns (APPLY_K, KDOT, DIFF_KT)           */
  int32_t na;          /* [in] DIFF_KT pair count (1..4) / KDOT ndot (1..2)      */
  /* [plan] scheduling, filled by lddmm_segments_plan */
  int64_t item0;       /* first pair-sum work item of this segment              */
  int64_t part0;       /* offset (doubles) of this segment's chunk partials     */
  int64_t fin0;        /* first finalize block of this segment                  */
  int64_t red0;        /* offset (doubles) of its per-block reduction partials  */
  int32_t row_tiles;   /* target tiles of LDDMM_ROWS_PER_ITEM rows              */
  int32_t chunks0;     /* source chunks over set 0                              */
  int32_t chunks1;     /* source chunks over set 1                              */
  int32_t chunk_len;   /* sources per chunk (multiple of the LDS tile)          */
  int32_t fin_blocks;  /* finalize blocks                                       */
  int32_t stage0;      /* first staged source tile: segments with a perm (set   *
                        * 0) and energy-only CURRENTS / MEASURE / KNORM sums     *
                        * (both sets) have their tiles staged once per launch in *
                        * the workspace and read by every row tile               */
  /* [in] optional processing order of FLOW_STEP, COVECTOR_STEP and KDOT segments (rows = sources),
   * NULL for the natural one: a permutation of 0..ny-1 in device memory, perm[i] =
   * the point taken as the i-th row and the i-th source.  Outputs stay indexed by
   * point, and the sums are the same up to rounding.  An order that keeps
   * neighbouring indices close in space makes the LDS tiles compact; the FP32
   * Gaussian sums then evaluate the exponent of compact 64-source blocks as
   * 2<x, y> - |y|^2 - |x|^2 in tile coordinates (3 instead of 6 instructions per
   * pair).  The permutation itself is not checked (it is not read on the host).
   * No counterpart in the reference, whose sums run in point order. */
  const int32_t *perm;
  /* [in] optional dispatch order of this segment's work items, NULL for item order:
   * a permutation of 0..n_order-1 in device memory, item_order[j] = the work item
   * (row tile, source chunk) that the segment's j-th workgroup runs.  n_order must be
   * the planned item count (row_tiles x (chunks0 + chunks1)); lddmm_segments_plan
   * rejects any other.  Every item writes its own partials, so the results are the
   * same bit for bit in any order; an order that starts with the items holding the
   * most unculled pairs (lddmm_item_order) shortens the launch's tail, where the
   * last workgroups run while the rest of the chip idles.  No counterpart in the
   * reference. */
  const int32_t *item_order;
  int64_t n_order;
} lddmm_seg_t;

/* Plan a segmented launch: fills the [plan] fields of segs[0..nseg) and
 * returns the number of pair-sum work items, finalize blocks and the
 * workspace size (descriptor head + chunk partials + reduction partials + the
 * staged source tiles of segments with a perm).
 * Host-only; touches no device memory.  Validates every segment (d, m, na,
 * kernel, order, sizes); all segments of one launch share the operation, the
 * kernel family, d, m and na. */
int lddmm_segments_plan(int op, int prec, int d, lddmm_seg_t *segs, int nseg,
                        int64_t *n_items, int64_t *n_fin_blocks,
                        size_t *work_bytes);

/* Run a planned segmented launch.  segs_host is the planned host array (its
 * first entry fixes the launch shape).  segs_dev is a device copy of it, or
 * NULL to have the library copy segs_host into the head of `work` on `stream`.
 * Results are deterministic: each row sums its sources in a fixed tile order,
 * chunk partials are combined in chunk order and totals are reduced in a fixed
 * tree; a segment's chunking depends only on its own sizes, so its result is
 * the same alone and inside any batch. */
int lddmm_segments_run(int op, int prec, int d, const lddmm_seg_t *segs_host,
                       const lddmm_seg_t *segs_dev, int nseg, int64_t n_items,
                       int64_t n_fin_blocks, void *work, size_t work_bytes,
                       void *stream);

/* n_steps launches in order on one stream: launch t runs the nseg descriptors
 * segs_host[t*nseg .. (t+1)*nseg) (device copy at segs_dev + t*nseg, required).
 * Every step must have the sizes of step 0, which was planned (n_items,
 * n_fin_blocks, work_bytes as returned for it) and whose scheduling fields
 * every step carries.  Used for the T time steps of Euler flows and adjoints,
 * where step t reads what step t-1 wrote.  Flow steps with staged tiles (segments
 * with a perm) reduce step t's totals in step t+1's staging launch and launch the
 * totals kernel once, after the last step; every total is written when the
 * call's work is complete, as with one call per step.  No counterpart in the
 * reference (its time loop calls the kernel sum once per step,
 * pointEvolution.py:691-697). */
int lddmm_segments_run_steps(int op, int prec, int d, const lddmm_seg_t *segs_host,
                             const lddmm_seg_t *segs_dev, int n_steps, int nseg,
                             int64_t n_items, int64_t n_fin_blocks, void *work,
                             size_t work_bytes, void *stream);

/* ---- one-call forms of the reference's f2py kernel sums ----------------
 * These plan one segment and copy its descriptor into the head of `work`. */

/* kernelFunctions.f90:1-54 applyk (via Kernel.applyK, kernelFunctions.py:310-369):
 * out(ny, m) = sum_l K(|y_k - x_l|) w(l, :).  y: (ny, d), x: (nx, d), w: (nx, m). */
size_t lddmm_apply_k_workspace_bytes(int64_t ny, int64_t nx, int d, int m);
int lddmm_apply_k(const double *y, const double *x, const double *w,
                  int64_t ny, int64_t nx, int d, int m, int kernel, int order,
                  double sigma, int prec, double *out, void *work,
                  size_t work_bytes, void *stream);

/* Allocation-free form of lddmm_apply_k.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory or pass a workspace (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernels and results as lddmm_apply_k, without (work, work_bytes): the
 * scratch is one fixed array of LDDMM_NOALLOC_SCRATCH_BYTES per GPU, in the
 * library's device data, held until the process ends; the call allocates nothing.
 * A request whose plan needs more scratch than that returns LDDMM_ESCRATCH before
 * launching anything.  96 MiB holds the plan of every square sum up to 4,000,000
 * points with m <= 3 at either precision (the chunk partials level off near
 * 92 MiB; 40,962 points need 51 MiB at FP32C, 31 MiB at FP64).
 * Every call on one GPU shares that GPU's array and leaves its contents behind:
 * calls that overlap on one GPU (different streams, or threads) race and must be
 * serialised by the caller.  `stream` must belong to the current device. */
#define LDDMM_NOALLOC_SCRATCH_BYTES ((size_t)96 << 20)
int lddmm_apply_k_noalloc(const double *y, const double *x, const double *w,
                          int64_t ny, int64_t nx, int d, int m, int kernel,
                          int order, double sigma, int prec, double *out,
                          void *stream);

/* kernelFunctions.f90:56-114 applykdifft (via Kernel.applyDiffKT,
 * kernelFunctions.py:413-449): a1 (na, ny, m), a2 (na, nx, m) -> out (ny, d). */
size_t lddmm_apply_diff_kt_workspace_bytes(int64_t ny, int64_t nx, int d,
                                           int m, int na);
int lddmm_apply_diff_kt(const double *y, const double *x, const double *a1,
                        const double *a2, int64_t ny, int64_t nx, int d, int m,
                        int na, int kernel, int order, double sigma, int prec,
                        double *out, void *work, size_t work_bytes,
                        void *stream);

/* kernelFunctions.f90:430-496 applykdiff11and12 (via
 * Kernel.applyDDiffK11and12, kernelFunctions.py:452-454):
 * x, a1, a2, p (n, d) -> out (n, d). */
size_t lddmm_apply_ddiff_k11and12_workspace_bytes(int64_t n, int d);
int lddmm_apply_ddiff_k11and12(const double *x, const double *a1,
                               const double *a2, const double *p, int64_t n,
                               int d, int kernel, int order, double sigma,
                               int prec, double *out, void *work,
                               size_t work_bytes, void *stream);

/* kernelFunctions.f90:499-550 applykmat (via Kernel.applyK(..., matrixWeights=True),
 * kernelFunctions.py:317-325): out(ny, m) = sum_l K(|y_k - x_l|) beta(k, l, :), beta
 * row-major (ny, nx, m), m = 1..3.  FP64 (the sum is bound by reading beta); Fortran
 * tables and cut-off; needs no workspace. */
int lddmm_apply_k_mat(const double *y, const double *x, const double *beta, int64_t ny,
                      int64_t nx, int d, int m, int kernel, int order, double sigma,
                      double *out, void *stream);

/* kernelFunctions.f90:552-610 applykdiffmat (via Kernel.applyDiffKmat,
 * kernelFunctions.py:405-411): out(ny, d) = sum_l 2K'(r^2) (y_k - x_l) beta(k, l),
 * beta row-major (ny, nx).  FP64; Laplacian orders 1..4. */
int lddmm_apply_diff_k_mat(const double *y, const double *x, const double *beta,
                           int64_t ny, int64_t nx, int d, int kernel, int order,
                           double sigma, double *out, void *stream);

/* Dense kernel matrix of Kernel.precompute (kernelFunctions.py:10-55,152-198,
 * 285-298): out(ny, nx) row-major = K(y_k, x_l) (deriv 0), K'(r^2) (deriv 1) or
 * K''(r^2) (deriv 2), FP64 with the dense path's exact-double Laplacian
 * polynomials and no cut-off; self_set (y is x, no firstVar / grid in the
 * reference) applies the reference's diagonal fills.  Writes 8 ny nx bytes. */
int lddmm_kernel_matrix(const double *y, const double *x, int64_t ny, int64_t nx, int d,
                        int kernel, int order, double sigma, int deriv, int self_set,
                        double *out, void *stream);

/* ---- geometry of the currents data term -------------------------------- */

/* surfaces.Surface.updateVertices (py-lddmm/surfaces.py:69-75):
 * centers(F,3) = (x_f0 + x_f1 + x_f2)/3, surfel(F,3) = (x_f1-x_f0) x (x_f2-x_f0).
 * faces: int32 (F, 3). */
int lddmm_surface_geometry(const double *vertices, const int32_t *faces,
                           int64_t nf, double *centers, double *surfel,
                           void *stream);

/* curves.Curve.updateVertices (py-lddmm/curves.py:67-73):
 * centers(E,d) = (x_e0 + x_e1)/2, lenel(E,d) = x_e1 - x_e0.  edges: int32 (E, 2). */
int lddmm_curve_geometry(const double *vertices, const int32_t *edges,
                         int64_t ne, int d, double *centers, double *lenel,
                         void *stream);

/* Vertex scatter of surfaces.currentNormGradient (surfaces.py:1069-1089) as a
 * deterministic gather: px[v] = scale * sum over the incidences (f, j) of v, in
 * the order given by the CSR (incid_ptr (nv+1), incid (3F) holding f*3+j), of
 *   dz[f] - (x_{f,j+2} - x_{f,j+1}) x z1[f].
 * With incidences sorted by (j, f) the sum order equals the reference's. */
int lddmm_surface_vertex_gradient(const double *vertices, const int32_t *faces,
                                  const double *z1, const double *dz,
                                  const int32_t *incid_ptr,
                                  const int32_t *incid, int64_t nv,
                                  double scale, double *px, void *stream);

/* Vertex scatter of curves.currentNormGradient (curves.py:604-613):
 * px[v] = scale * sum over incidences (e, j) of v of dz[e] + (2j-1) z1[e]. */
int lddmm_curve_vertex_gradient(const double *z1, const double *dz,
                                const int32_t *incid_ptr, const int32_t *incid,
                                int64_t nv, int d, double scale, double *px,
                                void *stream);

/* ---- batched geometry and gathers (no counterpart in the reference) -------
 * One launch updates the face geometry, or gathers the vertex gradients, of
 * many shapes at once: the serving layer's batched trials and gradients.  The
 * caller fills blk0 = first block of each entry (running sum of
 * ceil(count / 256) over the preceding entries) and passes the host array (for
 * the grid size) and a device copy (read by the kernel). */
typedef struct lddmm_geom {
  const double *vertices;  /* (nv, d) */
  const int32_t *faces;    /* (nf, k): k = 3 triangles (d = 3) or k = 2 edges */
  double *centers;         /* (nf, d) */
  double *normals;         /* (nf, d): surfel (triangles) or lenel (edges) */
  double *weights;         /* (nf) measure weights sqrt(|normal|^2 + 1e-10), or NULL */
  int64_t nf;
  int32_t k;
  int32_t d;
  int64_t blk0;
} lddmm_geom_t;

/* surfaces.Surface.updateVertices / curves.Curve.updateVertices for n shapes. */
int lddmm_geometry_batch(const lddmm_geom_t *geoms_host, const lddmm_geom_t *geoms_dev, int n,
                         void *stream);

typedef struct lddmm_gather {
  const double *vertices;  /* (nv, d); used by triangles */
  const int32_t *faces;    /* (nf, k) */
  const double *z1;        /* (nf, d) */
  const double *dz;        /* (nf, d) */
  const int32_t *incid_ptr;/* (nv + 1) */
  const int32_t *incid;    /* (k nf), entries f*k + j sorted by (vertex, corner, face) */
  double *px;              /* (nv, d) */
  int64_t nv;
  double scale;
  int32_t k;
  int32_t d;
  int64_t blk0;
} lddmm_gather_t;

/* lddmm_surface_vertex_gradient / lddmm_curve_vertex_gradient for n shapes. */
int lddmm_vertex_gradient_batch(const lddmm_gather_t *g_host, const lddmm_gather_t *g_dev, int n,
                                void *stream);

/* ---- landmark data term ---------------------------------------------------
 * Squared-distance data term of landmark-set matching, D(x) = sum_k |x_k - y_k|^2
 * over corresponding points (the energy of LandmarkObjective,
 * diffeoReg/landmarkMatching.h:49-62).  The gradient written is
 * grad = scale * (x - y); scale = 2 / sigma_err^2 gives dD/dx of D / sigma_err^2
 * (the reference's LandmarkObjective::gradient, landmarkMatching.h:63-69, uses
 * (x - y) / N instead, a rescaled direction for its own optimiser).
 * One workgroup per entry; the FP64 total follows one fixed order per entry,
 * so it is the same alone and in any batch. */
typedef struct lddmm_lmk {
  const double *x;  /* (n) deformed point coordinates, n = points * d */
  const double *y;  /* (n) target coordinates */
  double *total;    /* 1 double: sum (x - y)^2, or NULL */
  double *grad;     /* (n): scale * (x - y), or NULL */
  int64_t n;
  double scale;
} lddmm_lmk_t;

int lddmm_landmark_batch(const lddmm_lmk_t *host, const lddmm_lmk_t *dev, int n, void *stream);

/* ---- processing order (host only) ----------------------------------------
 * No counterpart in the reference (its sums run in point order).  They build
 * and check the optional lddmm_seg_t.perm of flow steps and RKHS products. */

/* out[i][k] = a[i][k] + (s[i] * d[i][k]) for k < len[i], i < n (device FP64 arrays; the
 * product rounded before the sum, as NumPy's a + s * d and two elementwise passes
 * round it; out may be a).  The line search's trial momenta of a batch of
 * registrations, at - eps dir (surfaceMatching.py:255-289), in one launch per 32
 * items.  Host arrays of pointers and sizes; enqueued on `stream`. */
int lddmm_axpy_batch(int n, double *const *out, const double *const *a, const double *const *d,
                     const double *s, const int64_t *len, void *stream);

/* Host geometry of a triangulated surface (row-major (nv, 3) vertices, (nf, 3)
 * int64 faces; a negative index counts from the end): face centers
 * ((x1 + x2) + x3) / 3, surfels (x2 - x1) x (x3 - x1) and the triple products
 * x1 . (x2 x x3), each in the reference's NumPy evaluation order and without
 * contraction, so bit for bit the values of surfaces.py:62-70 (computeCentersAreas)
 * and of the signed volume's terms (surfaces.py:711-717).  Any output may be NULL.
 * Replaces the NumPy expressions on the host side of Surface construction. */
int lddmm_surface_geometry_host(const double *v, int64_t nv, const int64_t *f, int64_t nf,
                                double *centers, double *surfel, double *det);

/* Spatial processing order of n points (row-major (n, d) host array, d = 1..3):
 * recursive bisection along the longest side of the bounding box, with split
 * points at multiples of 256 (then 64), so that every 256-point tile and every
 * 64-point quarter of the order is one compact subtree.  Writes order[i] = the
 * point taken i-th (a permutation of 0..n-1); deterministic. */
int lddmm_spatial_order(const double *points, int64_t n, int d, int32_t *order);

/* Fraction of the 64-point quarters of `order` whose points all lie within
 * squared distance r2max of their tile origin (the mean of the tile's points
 * 32, 96, 160, 224, clamped to the tile), distances taken after scaling the
 * coordinates by `scale`: the quarters an FP32 Gaussian pair kernel evaluates
 * in dot-product form (scale = sqrt(log2(e) / 2) / sigma, r2max = 2).
 * Returns -1 for invalid arguments. */
double lddmm_order_compactness(const double *points, int64_t n, int d, const int32_t *order,
                               double scale, double r2max);

/* Fraction of the (row block, 64-point quarter) blocks of a self sum over the
 * points in `order` whose bounding boxes lie within squared distance cut2 of each
 * other: the blocks a pair kernel evaluates rather than culls.  A row block is
 * `rows` consecutive points of the order (a wave's rows: 128 in FP32C, 64 in
 * FP64; a multiple of 64), rows past the end repeating the last point.  Beyond
 * max_pairs block pairs every k-th row block is taken.  Host code; returns -1 for
 * invalid arguments. */
double lddmm_order_live_fraction(const double *points, int64_t n, int d, const int32_t *order,
                                 int rows, double cut2, int64_t max_pairs);

/* Dispatch order of a segment's work items, longest first (the seg_t field
 * item_order).  seg: the segment as it will be launched (its sizes, kernel, sigma,
 * perm and output flags; device pointers may still be NULL, they are not read);
 * y: host copy of its rows (ny, d), which are also its set-0 sources; x1: host copy
 * of its set-1 sources (nx1, d) for CURRENTS and MEASURE segments with nx1 > 0,
 * else NULL; perm: host copy of its processing order (ny), NULL when it has none.
 * Plans the segment, estimates each work item's cost from bounding boxes (the
 * (wave, 64-source quarter) blocks that survive the pair kernel's cull, the
 * symmetric energy sums' skipped tiles, FP64 energies' near blocks counted double)
 * and writes into order[0..n) the items sorted by decreasing cost, ties in item
 * order; n must be the planned item count.  With order NULL nothing is written and
 * the planned item count is returned (y, x1, perm, n unused).  Host-only,
 * deterministic.  Supported for FLOW_STEP, COVECTOR_STEP, KDOT, KNORM, CURRENTS
 * and MEASURE (others return LDDMM_EUNSUPPORTED).  The geometry only steers the schedule: an order estimated
 * from earlier positions (a registration's template, reused for every step and
 * trial) leaves every result unchanged.  No counterpart in the reference. */
int lddmm_item_order(int op, int prec, int d, const lddmm_seg_t *seg, const double *y,
                     const double *x1, const int32_t *perm, int32_t *order, int64_t n);

/* ---- chunking targets (no counterpart in the reference: its sums are one
 * OpenMP loop over rows, kernelFunctions.f90:1-54) ------------------------ */

/* lddmm_segments_plan cuts a segment of ny rows into chunks of source tiles so
 * that (row tiles) x (chunks) is about a target number of work items.  The
 * target is kept per operation, precision, energy-only form (a data-term sum
 * with no per-row output) and size class of ny; items <= 0 restores the
 * built-in value (6,144; 24,576 for energy-only sums).  A setting applies to
 * plans made after it, in this process; set targets before planning, not
 * concurrently with it.  A segment's chunking still depends only on its own
 * sizes, so results stay the same alone and in any batch.
 * lddmm_apply_k_noalloc's scratch bound holds for the built-in targets of
 * LDDMM_OP_APPLY_K; a larger one can make it return LDDMM_ESCRATCH. */
#define LDDMM_SIZE_CLASSES 6
/* Size class of a segment of ny rows: 0 below 1,024 rows, 1 below 4,096, 2
 * below 16,384, 3 below 65,536, 4 below 262,144, 5 from 262,144 on. */
int lddmm_size_class(int64_t ny);
int lddmm_set_items_target(int op, int prec, int energy_only, int size_class, int64_t items);
/* The target in force (built-in when unset), or LDDMM_EINVAL. */
int64_t lddmm_get_items_target(int op, int prec, int energy_only, int size_class);

/* ---- template decimation ------------------------------------------------ */

/* surfaces.Surface.Simplify (py-lddmm/surfaces.py:244-259, used by SurfaceMatching's
 * subsampleTargetSize, surfaceMatching.py:131-133): decimate a triangle mesh to at most
 * `target` vertices by quadric-error edge collapse (Garland & Heckbert 1997; the
 * reference calls VTK's vtkQuadricDecimation, which implements the same method).  Host
 * code, no device.  out_vertices holds nv x 3 and out_faces nf x 3 entries (the result is
 * never larger); *nv_out / *nf_out receive the sizes kept.  Collapses that would make the
 * mesh non-manifold or fold a face are skipped, so a mesh may stop above `target`.
 * Deterministic.  LDDMM_EINVAL for face indices out of range or non-finite vertices. */
int lddmm_decimate_surface(const double *vertices, int64_t nv, const int32_t *faces,
                           int64_t nf, int64_t target, double *out_vertices,
                           int32_t *out_faces, int64_t *nv_out, int64_t *nf_out);

/* ---- small helpers ------------------------------------------------------ */

/* Deterministic FP64 sum of x[0..n) into *out (single fixed reduction tree). */
int lddmm_sum_f64(const double *x, int64_t n, double *out, void *stream);

/* Last HIP error code seen by the library (0 if none). */
int lddmm_last_hip_error(void);
/* Library version string. */
const char *lddmm_version(void);

#ifdef __cplusplus
}
#endif

#endif /* LDDMM_H */
