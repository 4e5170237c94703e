/*
 * flof.h -- C-ABI of the MI355X FlOF 4D deformation library (libflof_hip.so).
 *
 * The entry points replace the mantaflow plugin surface that scenes/flof.py
 * and the 4D tests call.  Each declaration cites the reference function it
 * stands in for (paths relative to the ofblend source tree).  Plain pointers,
 * sizes and PODs only: no C++ or torch types cross this boundary, no exception
 * escapes it.  Every function returns FLOF_OK (0) or a negative status; the
 * message of the last failure is available from flof_last_error().
 *
 * Threading: one flof_ctx per host thread.  All work is enqueued on the
 * context's HIP stream.  Functions that return a host value (reductions,
 * solve statistics, downloads) synchronise the stream before returning, the
 * others return after enqueueing (call flof_sync() to wait), mirroring the
 * synchronous plugin calls of the reference when the shim syncs.
 */
#ifndef FLOF_H
#define FLOF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum flof_status {
    FLOF_OK = 0,
    FLOF_ERR_ARG = -1,       /* bad argument / null pointer                  */
    FLOF_ERR_DIM = -2,       /* grid dimensions mismatch or too small        */
    FLOF_ERR_TYPE = -3,      /* element type not supported by the function   */
    FLOF_ERR_HIP = -4,       /* HIP runtime error                            */
    FLOF_ERR_NOMEM = -5,     /* device allocation failed                     */
    FLOF_ERR_STATE = -6,     /* solver state invalid (e.g. dt != 1)          */
    FLOF_ERR_RANGE = -7      /* value outside the supported range            */
};

/* Element types; values equal Grid4dBase::Grid4dType (source/grid4d.h:29-35). */
enum flof_elem {
    FLOF_REAL = 1,
    FLOF_INT = 2,
    FLOF_VEC4 = 8
};

/* Storage precision of a grid.  F32 is the reference's `Real`
 * (source/util/vectorbase.h:42-51); BF16/F16 are the mixed-precision storage
 * variants (arithmetic always accumulates in FP32, CG scalars in FP64). */
enum flof_dtype {
    FLOF_F32 = 0,
    FLOF_BF16 = 1,
    FLOF_F16 = 2
};

typedef struct flof_ctx flof_ctx;
typedef struct flof_grid4d flof_grid4d;
typedef struct flof_warp_state flof_warp_state;
typedef struct flof_warmnet flof_warmnet;

/* Parameters of opticalFlowMultiscale4d / opticalFlow4d.  Field names and
 * defaults follow the keyword arguments of
 * source/plugin/optflow4d.cpp:2182-2189 (multiscale) and :2110-2114 (single
 * level).  Fill with flof_of_params_default(). */
typedef struct flof_of_params {
    double wSmooth;        /* 0.     smoothness weight                        */
    double wEnergy;        /* 0.     Tikhonov weight                          */
    int32_t level;         /* 0      recursion level of the call              */
    double postVelBlur;    /* 0.     Gaussian blur of each solve (sigma*2)    */
    double cgAccuracy;     /* 1e-4   relative signed-max residual target      */
    int32_t blurType;      /* 1      only 1 (Gaussian) is defined             */
    double cfl;            /* 999    advection CFL                            */
    int32_t orderTime;     /* 1      unused by the reference                  */
    int32_t orderSpace;    /* 1      unused by the reference                  */
    double resetBndWidth;  /* -1.    zero band after each solve (fraction)    */
    int32_t multiStep;     /* 1      OF solves per level                      */
    int32_t projSizeThresh;/* 9999   projection-only solve above this size    */
    int32_t minGridSize;   /* 10     recursion stops at X <= minGridSize      */
    int32_t doFinalProject;/* 0      level-set projection at level 0          */
    int32_t cgMaxIter;     /* 1000   CG iteration cap (optflow4d.cpp:502)     */
    /* Extensions of this library (no reference counterpart): */
    int32_t warmStart;     /* 0      1: use *warm as CG start at each level   */
    int32_t precision;     /* 0      storage precisions, a combination of:
                                     FLOF_PSTORE_BF16: the CG search
                                     direction in bfloat16;
                                     FLOF_PREC_IMG_BF16: the images (i0, i1,
                                     their pyramid levels and warps) in
                                     bfloat16 and the gradients of i1 rounded
                                     to binary16 (the mixed-precision solve;
                                     no projection, rhsT, network or
                                     capture).  0 (FLOF_PSTORE_F32) is the
                                     reference's FP32 arithmetic; stencil
                                     arithmetic is FP32 and CG scalars FP64
                                     in every form                            */
} flof_of_params;

enum flof_pstore { FLOF_PSTORE_F32 = 0, FLOF_PSTORE_BF16 = 1, FLOF_PREC_IMG_BF16 = 2 };

#define FLOF_MAX_LEVELS 16
#define FLOF_MAX_STEPS 10

/* Convergence record of one call; level 0 is the finest level. */
typedef struct flof_solve_info {
    int32_t levels;                              /* pyramid levels used      */
    int32_t dims[FLOF_MAX_LEVELS][4];            /* X,Y,Z,T per level        */
    int32_t steps[FLOF_MAX_LEVELS];              /* OF solves run per level  */
    int32_t iters[FLOF_MAX_LEVELS][FLOF_MAX_STEPS];  /* CG iterations        */
    double residual[FLOF_MAX_LEVELS][FLOF_MAX_STEPS];/* final signed-max
                                                    residual / initial one  */
    int32_t status[FLOF_MAX_LEVELS][FLOF_MAX_STEPS]; /* 0 converged, 1 not
                                                    converged, 2 NaN/zero
                                                    sigma, 3 rhs below eps  */
    double err_level[FLOF_MAX_LEVELS];           /* calcLsDiff after the
                                                    level's pre-warp         */
    double err_pre;                              /* error before any solve   */
    double err_final;                            /* error after final warp   */
    int32_t warm_used[FLOF_MAX_LEVELS];          /* 1: warm start kept, 0:
                                                    cold start (or fallback) */
    int32_t pad_;
    int64_t total_iters;
    double residual_l2[FLOF_MAX_LEVELS][FLOF_MAX_STEPS]; /* ||r||_2 / ||b||_2
                                                    of the final residual (an
                                                    addition; b: the rhs)    */
} flof_solve_info;

/* ---- context ----------------------------------------------------------- */
void flof_of_params_default(flof_of_params *p);
int flof_ctx_create(flof_ctx **out, int device);
int flof_ctx_destroy(flof_ctx *ctx);
/* Message of the last failure of ctx; with ctx == NULL, of the last failed
 * call without a context (flof_calc_ls_diff4d_noalloc) on this thread. */
const char *flof_last_error(const flof_ctx *ctx);
int flof_sync(flof_ctx *ctx);
/* HIP stream (hipStream_t) the context enqueues on. */
void *flof_ctx_stream(flof_ctx *ctx);
/* Release cached temporaries of the context's device pool. */
int flof_ctx_trim(flof_ctx *ctx);
/* Per-kernel timing with HIP events on the context stream.  While enabled,
 * every launch of kernel class `kind` (FLOF_KT_*) is bracketed by events;
 * flof_ctx_timing returns the number of launches and their summed duration. */
enum flof_kernel_kind {
    FLOF_KT_CG_APPLY = 0,   /* p.(A p)                                 */
    FLOF_KT_CG_UPDATE = 1,  /* r update, signed max of r, z.r         */
    FLOF_KT_CG_DIR = 2,     /* x update, p = z + beta p                */
    FLOF_KT_WARP = 3,       /* 3D frame warp                           */
    FLOF_KT_ADVECT = 4,     /* 4D semi-Lagrangian advection            */
    FLOF_KT_COUNT = 8
};
int flof_ctx_enable_timing(flof_ctx *ctx, int kind, int enable);
int flof_ctx_timing(flof_ctx *ctx, int kind, int64_t *launches, double *ms_total);
/* The same for the launches of one grid size only, without clearing the
 * records (flof_ctx_timing clears them): PCG launches are tagged with the
 * cell count of their level (X*Y*Z*T); tag < 0 takes all. */
int flof_ctx_timing_tag(flof_ctx *ctx, int kind, int64_t tag, int64_t *launches,
                        double *ms_total);

/* Autotuner of the PCG kernels' launch shape (no reference counterpart):
 * while enabled, the first solve of each grid shape times one PCG iteration
 * for each candidate -- the t-marching kernels at 2, 4, 8 and 16 workgroups
 * per CU, and, where the LDS-tiled stencil kernels apply (FP32 images, the
 * stencil choice not forced to the t-marching kernels, grids below 2^31
 * cells), their 64x4x4-cell tiles with t segments of the lengths that 1, 2,
 * 4 and 8 work items per CU give -- and keeps the fastest for that shape
 * (results do not depend on it).  flof_ctx_tuned_blocks returns the
 * t-marching workgroups per CU for a shape, or 0 if it has not been tuned;
 * flof_ctx_tuned_shape returns 1 if not tuned, else FLOF_OK and the choice:
 * t-marching workgroups per CU, LDS work items per CU (0: the t-marching
 * stencil kernels) and the LDS tiles' t-segment length (0 likewise). */
int flof_ctx_set_autotune(flof_ctx *ctx, int enable);
int flof_ctx_tuned_blocks(const flof_ctx *ctx, int X, int Y, int Z, int T);
int flof_ctx_tuned_shape(const flof_ctx *ctx, int X, int Y, int Z, int T, int *tm_bpc,
                         int *lds_bpc, int *lds_seglen);

/* Kernels of the PCG's two stencil passes, p.(A p) and the update (no
 * reference counterpart; results are identical): FLOF_STENCIL_LDS tiles
 * 64x4x4 cells in LDS, FLOF_STENCIL_TMARCH reads the neighbours through the
 * caches, FLOF_STENCIL_AUTO (the default) takes the LDS tiles for levels
 * where they are faster (FP32 direction storage, t segments of 16 or more
 * steps, X filling 90% of the 64-wide tiles) and the t-marching kernels
 * elsewhere; FLOF_STENCIL_QARRAY is the first GPU form (three streaming
 * kernels that store q = A p and g, 208 B per cell and iteration, FP32
 * direction storage only), kept for comparison. */
enum flof_stencil {
    FLOF_STENCIL_AUTO = 0,
    FLOF_STENCIL_TMARCH = 1,
    FLOF_STENCIL_LDS = 2,
    FLOF_STENCIL_QARRAY = 3
};
int flof_ctx_set_stencil(flof_ctx *ctx, int kind);

/* ---- level-set mesh of a 3D frame and its .bobj.gz file ---------------- */
typedef struct flof_mesh flof_mesh;
/* LevelsetGrid::createMesh (source/levelset.cpp:266-365) of a 3D level set
 * phi (a float[Z][Y][X] device buffer, X, Y, Z >= 2): marching cubes on the
 * iso value 1e-4 of -phi, cubes with a value at or below -1000 skipped; the
 * nodes and triangles are numbered as the reference numbers them.  The
 * mesh's normals are those its file holds (Mesh::computeVertexNormals,
 * source/mesh.cpp:460-479); the mesher's own gradient normals, which
 * writeBobjFile replaces, are not formed.  Grids of 2^31/5 points or more
 * are rejected (the counts are 32-bit). */
int flof_create_mesh(flof_ctx *ctx, const float *phi_dev, int X, int Y, int Z,
                     flof_mesh **out);
int flof_mesh_size(const flof_mesh *m, int64_t *nodes, int64_t *tris);
/* Host copies: pos and normals 3 floats per node, tris 3 ints per triangle
 * (any pointer may be NULL). */
int flof_mesh_download(const flof_mesh *m, float *pos, float *normals, int32_t *tris);
/* writeBobjFile (source/fileio.cpp:121-160): gzip, int nodes, positions
 * (p - gs/2) / max(gs) for the level set's grid size gs, int nodes,
 * normals, int triangles, 3 ints each. */
int flof_mesh_save_bobj(const flof_mesh *m, const char *name);
int flof_mesh_destroy(flof_mesh *m);

/* ---- Output post-processing of warped 3D frames (scenes/flof.py:737-782,
 * scenes/ofHelpers.py:108-115).  A frame is a float[Z][Y][X] device buffer;
 * Z == 1 is a 2D grid. */

/* Scalar frame operations the union blends and the temporal merge are made
 * of: COPY a = b (Grid::copyFrom, source/grid.cpp:198-207), SET_CONST a = s
 * (grid.cpp:230-233), MULT_CONST a *= s (gridMultScalar, source/grid.h:735),
 * ADD_SCALED a += s * b (gridScaledAdd, grid.h:740-743), JOIN a = min(a, b)
 * (the level-set union KnJoin, source/levelset.cpp:114-118).  n cells. */
enum flof_frame_op_kind {
    FLOF_FRAME_COPY = 0,
    FLOF_FRAME_SET_CONST = 1,
    FLOF_FRAME_MULT_CONST = 2,
    FLOF_FRAME_ADD_SCALED = 3,
    FLOF_FRAME_JOIN = 4
};
int flof_frame_op(flof_ctx *ctx, int op, float *a, const float *b, int64_t n, float s);

/* simpleBlurSpecial (source/test.cpp:94-127): iter passes of the 6- (2D: 4-)
 * neighbour mean over the cells >= thresh inside the bord band, the outer
 * layer of the grid swapped with a zeroed temporary like the reference's. */
int flof_simple_blur_special(flof_ctx *ctx, float *a, int X, int Y, int Z, int iter,
                             float thresh, int bord);

/* Grid::setBound (source/grid.cpp:462-473): value into the cells within
 * width of the border (z only for 3D grids). */
int flof_frame_set_bound(flof_ctx *ctx, float *a, int X, int Y, int Z, float value,
                         int width);

/* Measurement aid (no reference counterpart): the bandwidth a plain
 * streaming kernel reaches with the byte mix of one PCG kernel and no
 * stencil work, on `cells` cells of scratch memory taken from the context's
 * pool: FLOF_MIX_UPDATE reads 3 float4 + 1 float and writes 2 float4 per
 * cell (84 B: an update that also moves x), FLOF_MIX_DIR reads 2 float4 +
 * 1 float and writes 1 float4 (52 B: the update and the direction kernel of
 * even iterations), FLOF_MIX_COPY reads and writes 1 float4 (32 B),
 * FLOF_MIX_DIR_X reads 4 float4 + 1 float and writes 2 float4 (100 B: the
 * direction kernel of odd iterations, which also moves x); non-temporal
 * accesses, best of 1, 2, 4, 8 and 16 workgroups of 256 threads per CU,
 * each timed over `reps` launches.  *gbps = bytes moved / time, in GB/s. */
enum flof_mix {
    FLOF_MIX_UPDATE = 0,
    FLOF_MIX_DIR = 1,
    FLOF_MIX_COPY = 2,
    FLOF_MIX_DIR_X = 3
};
int flof_stream_bandwidth(flof_ctx *ctx, int mix, int64_t cells, int reps, double *gbps);

/* Verification aid (no reference counterpart): compares the fast reciprocal
 * the PCG kernels use for the Jacobi preconditioner (taken on a level only
 * when every diagonal entry lies in [2^-126, 2^126)) with the IEEE division
 * 1.f / d for every float whose bit pattern lies in [lo, hi] and stores the
 * number of differing results (bitwise) in *mismatches.  The preconditioner
 * must be the correctly rounded 1/d that the reference stores. */
int flof_check_reciprocal(flof_ctx *ctx, uint32_t lo, uint32_t hi, uint64_t *mismatches);

/* ---- Grid4d storage (source/grid4d.h:117-265, source/grid4d.cpp:60-138) - */
/* Creates a zero-initialised X*Y*Z*T grid (x fastest, then y, z, t). */
int flof_grid4d_create(flof_ctx *ctx, int X, int Y, int Z, int T, int elem,
                       int dtype, flof_grid4d **out);
int flof_grid4d_destroy(flof_grid4d *g);
int flof_grid4d_dims(const flof_grid4d *g, int dims[4], int *elem, int *dtype);
/* Device address of the grid's storage (for zero-copy interop). */
void *flof_grid4d_device_ptr(flof_grid4d *g);
/* Host <-> device copies of the whole grid; `bytes` must equal its size. */
int flof_grid4d_upload(flof_grid4d *g, const void *host, size_t bytes);
int flof_grid4d_download(const flof_grid4d *g, void *host, size_t bytes);
/* Grid4d::clear / copyFrom / swap (grid4d.cpp:99-122,198-207). */
int flof_grid4d_clear(flof_grid4d *g);
int flof_grid4d_copy_from(flof_grid4d *dst, const flof_grid4d *src);
int flof_grid4d_swap(flof_grid4d *a, flof_grid4d *b);
/* Elementwise ops (grid4d.cpp:208-245).  `value` points to 1 float for REAL
 * and INT (converted) grids, 4 floats for VEC4 grids. */
int flof_grid4d_set_const(flof_grid4d *g, const float *value);
int flof_grid4d_add_const(flof_grid4d *g, const float *value);
int flof_grid4d_mult_const(flof_grid4d *g, const float *value);
int flof_grid4d_add(flof_grid4d *g, const flof_grid4d *a);
int flof_grid4d_sub(flof_grid4d *g, const flof_grid4d *a);
int flof_grid4d_mult(flof_grid4d *g, const flof_grid4d *a);
int flof_grid4d_add_scaled(flof_grid4d *g, const flof_grid4d *a,
                           const float *factor);
int flof_grid4d_clamp(flof_grid4d *g, float vmin, float vmax);
/* getMax / getMin / getMaxAbs (grid4d.cpp:247-270): for VEC4 grids the
 * square root of the max/min squared norm. */
int flof_grid4d_get_max(flof_grid4d *g, float *out);
int flof_grid4d_get_min(flof_grid4d *g, float *out);
int flof_grid4d_get_max_abs(flof_grid4d *g, float *out);
/* setBound / setBoundNeumann (grid4d.cpp:355-412). */
int flof_grid4d_set_bound(flof_grid4d *g, const float *value, int width);
int flof_grid4d_set_bound_neumann(flof_grid4d *g, int width);
/* grid4dMaxDiff / grid4dMaxDiffVec4 (grid4d.cpp:419-464); *out holds the
 * reference's Real (FP32) result. */
int flof_grid4d_max_diff(const flof_grid4d *a, const flof_grid4d *b,
                         double *out);
/* setRegion4d / setRegion4dVec4 (grid4d.cpp:466-483). */
int flof_set_region4d(flof_grid4d *dst, const float start[4],
                      const float end[4], const float *value);
/* getSliceFrom4d (grid4d.cpp:486-499): dst is a host-visible-size 3D device
 * buffer of dx*dy*dz floats. */
int flof_get_slice_from4d(flof_grid4d *src, int srct, float *dst3d_dev,
                          int dx, int dy, int dz);
/* placeGrid3d (grid4d.cpp:516-524). */
int flof_place_grid3d(const float *src3d_dev, int sx, int sy, int sz,
                      flof_grid4d *dst, int dstt);

/* ---- resampling (grid4d.cpp:531-569, optflow4d.cpp:40-57) -------------- */
/* interpolateGrid4d / interpolateGrid4dVec: target(c) = interp(source,
 * c*f + off) with f,off from gridFactor4d(offset, scale, size). */
int flof_interpolate_grid4d(flof_grid4d *target, flof_grid4d *source,
                            const float offset[4], const float scale[4],
                            const float size[4]);

/* ---- advection (optflow4d.cpp:1275-1321,2170-2180) --------------------- */
/* advect4d(vel, grid, dtFac) with solver timestep `dt`: semi-Lagrangian
 * quadrilinear backtrace of `grid` (REAL or VEC4) by dt*dtFac*vel. */
int flof_advect4d(flof_grid4d *vel, flof_grid4d *grid, float dt, float dtFac);
/* advectCflTemplate<Grid4d>: CFL-substepped advection by velFactor*vel. */
int flof_advect_cfl4d(flof_grid4d *vel, flof_grid4d *grid, float cfl,
                      float velFactor, float dt);

/* ---- error metric (optflow4d.cpp:895-927,2132-2137) -------------------- */
int flof_calc_ls_diff4d(flof_grid4d *i0, flof_grid4d *i1, flof_grid4d *out,
                        float correction, int bnd, float *result);

/* calcLsDiff4d without any allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The same
 * kernels as flof_calc_ls_diff4d on plain device pointers (grids of X*Y*Z*T
 * floats, x fastest; out may be NULL), enqueued on `stream` (a hipStream_t
 * of the current device, NULL: the null stream) without a host sync: the
 * reference's Real result is stored to the device float *result_dev.  The
 * workgroup partials go to a fixed per-device scratch array of 2048
 * double-double partials (32 KiB, held by each GPU until the process ends),
 * so calls that overlap on one GPU (on different streams) race.  Grids of
 * 2^31 cells or more are rejected (FLOF_ERR_RANGE); errors are reported by
 * flof_last_error(NULL) on the calling thread. */
int flof_calc_ls_diff4d_noalloc(const float *i0, const float *i1, float *out, int X,
                                int Y, int Z, int T, float correction, int bnd,
                                void *stream, float *result_dev);

/* ---- blur (optflow4d.cpp:127-174) -------------------------------------- */
int flof_gaussian_blur4d(flof_grid4d *grid, float sigma, int iter);

/* ---- optical flow (optflow4d.cpp:361-553,2110-2119) -------------------- */
/* opticalFlow4d: one level.  vel (VEC4) holds the incoming flow used by the
 * right-hand side and receives the solution.  rhsT (REAL, may be NULL)
 * receives component 0 of the right-hand side. */
int flof_optical_flow4d(flof_grid4d *vel, flof_grid4d *i0, flof_grid4d *i1,
                        flof_grid4d *rhsT, const flof_of_params *p,
                        flof_solve_info *info);
/* opticalFlowMultiscale4d (optflow4d.cpp:936-1173,2182-2195).  Returns the
 * level-0 error through *err (the reference's return value).  When
 * p->warmStart is set, `warm` (VEC4, finest size) is an initial guess of the
 * finest-level CG (restricted for coarser levels); the converged flow is the
 * one of the same linear systems. */
int flof_optical_flow_multiscale4d(flof_grid4d *vel, flof_grid4d *i0,
                                   flof_grid4d *i1, flof_grid4d *rhsT,
                                   const flof_of_params *p,
                                   flof_grid4d *warm, flof_solve_info *info,
                                   float *err);
/* opticalFlowMultiscale4d with the learned warm start (no reference
 * counterpart: the reference starts every level's CG from zero,
 * optflow4d.cpp:284): on each of the `levels` finest pyramid levels (-1: all)
 * that run one plain solve, the network `net` (of vel's context) predicts
 * the level's increment from the level's pre-warped i0 and its i1, and the
 * CG starts from that guess scaled energy-optimally, or from zero where the
 * scaled guess has a larger residual norm than zero (the fallback).  The
 * converged flow solves the same linear systems as a cold start. */
int flof_optical_flow_multiscale4d_net(flof_grid4d *vel, flof_grid4d *i0, flof_grid4d *i1,
                                       const flof_of_params *p, flof_warmnet *net,
                                       int levels, flof_solve_info *info, float *err);
/* opticalFlowMultiscale4d that also copies out the pre-warped i0 of pyramid
 * level `level` (0: the finest) and the increment that level's solve adds to
 * the flow (the training pairs of the warm-start network; no reference
 * counterpart).  i0w_out (REAL) and inc_out (VEC4) have that level's size;
 * multiStep must be 1; vel receives the same flow as without the copies. */
int flof_optical_flow_multiscale4d_capture(flof_grid4d *vel, flof_grid4d *i0,
                                           flof_grid4d *i1, const flof_of_params *p,
                                           int level, flof_grid4d *i0w_out,
                                           flof_grid4d *inc_out, flof_solve_info *info,
                                           float *err);
/* Batched opticalFlowMultiscale4d (no reference counterpart; the reference
 * runs its solves one after another): n independent solves, sizes may
 * differ, all with parameters *p (warmStart must be 0), run on
 * `concurrency` worker streams of ctx's device (<= 0: 4, at most 16).
 * Each result is identical to a flof_optical_flow_multiscale4d call.
 * infos and errs may be null, else arrays of n. */
int flof_optical_flow_multiscale4d_batch(flof_ctx *ctx, int n,
                                         flof_grid4d *const *vel,
                                         flof_grid4d *const *i0,
                                         flof_grid4d *const *i1,
                                         const flof_of_params *p,
                                         int concurrency,
                                         flof_solve_info *infos, float *errs);
/* Fused batch of opticalFlowMultiscale4d (no reference counterpart; the
 * reference solves one problem at a time, optflow4d.cpp:274-329 per level):
 * n independent solves, sizes may differ, all with parameters *p, which
 * must ask for one plain solve per level (multiStep 1, postVelBlur 0,
 * resetBndWidth <= 0, no final projection, no warm start).  The pyramids
 * are built first; then the levels of equal index of all problems are
 * solved together, every PCG iteration as one launch per kernel kind that
 * spans all of them (a work-item table maps each workgroup to its problem
 * and its solver state; converged problems return at once), chunks of
 * iterations captured as HIP graphs.  A level of 2^22 cells or more whose
 * single solve would run the LDS-tiled stencil kernels iterates on its own
 * instead (those kernels run one problem per launch; the level fills the
 * chip by itself).  Runs on ctx's stream; each result is
 * bit-identical to a flof_optical_flow_multiscale4d call.  The finest levels
 * of all problems are resident at once (96 B of PCG workspace per cell). */
int flof_optical_flow_multiscale4d_fused(flof_ctx *ctx, int n, flof_grid4d *const *vel,
                                         flof_grid4d *const *i0, flof_grid4d *const *i1,
                                         const flof_of_params *p, flof_solve_info *infos,
                                         float *errs);
/* corrVelsOf4d (optflow4d.cpp:651-802,2121-2130). */
int flof_corr_vels_of4d(flof_grid4d *dst, flof_grid4d *vel,
                        flof_grid4d *phiOrg, flof_grid4d *phiCurr,
                        flof_grid4d *phiTarget, float threshPhi,
                        float threshNorm, float postVelBlur,
                        float resetBndWidth, int maxIter);

/* ---- full-resolution frames (optflow4d.cpp:1445-1595,1622-2105) -------- */
/* loadPlaceGrid4d with the frame files already decoded into device memory:
 * `frames_dev` holds (fileIdxEnd - fileIdxStart) frames of sx*sy*sz floats,
 * frame f at offset f*sx*sy*sz (file index fileIdxStart + f). */
int flof_load_place_grid4d(const float *frames_dev, int sx, int sy, int sz,
                           flof_grid4d *phi, const float offset[4],
                           const float scale[4], int fileIdxStart,
                           int fileIdxEnd, int debugSkipLoad, float spread,
                           const float overrideSize[4], float overrideTimeOff,
                           int overrideGoodRegion, float loadTimeScale,
                           int rescaleSdfValues, float sdfIsoOff,
                           float repeatStartFrame);
/* loadAdvectTimeSlice_OptRun for one deformation volume (useDefoVols off):
 * the two deformation time slices are read from the device-resident 4D
 * deformation `defo` instead of a .uni file.  dst is a dx*dy*dz device
 * buffer; cells outside the bnd=10 / bordSkip band are left unchanged. */
int flof_advect_time_slice_opt(flof_grid4d *defo, float *dst_dev, int dx,
                               int dy, int dz, flof_grid4d *phi, float time,
                               float blendAlpha, float loadTimeScale,
                               const float defoOffset[4],
                               const float defoScale[4],
                               const float defoFactor[4],
                               const float overrideSize[4],
                               float overrideTimeOff, int zeroVel,
                               int bordSkip, float defoAniFac);
/* loadAdvectTimeSlice (optflow4d.cpp:1671-1760), the non-cached variant:
 * the same lookup over the whole bnd=1 interior of dst (no border skip). */
int flof_advect_time_slice(flof_grid4d *defo, float *dst_dev, int dx, int dy,
                           int dz, flof_grid4d *phi, float time,
                           float blendAlpha, float loadTimeScale,
                           const float defoOffset[4], const float defoScale[4],
                           const float defoFactor[4],
                           const float overrideSize[4], float overrideTimeOff,
                           int zeroVel, float defoAniFac);
/* loadAdvectTimeSlice_OptRun with deformation volumes (useDefoVols;
 * optflow4d.cpp:1786-1870, 2013-2085): 2 or 3 VEC4 volumes of one size,
 * each seen through a window of int(T * max(0.2, partialLoadFac)) slices
 * around the time; 2 volumes: vt = alpha dv1(p - dv2(p)) + thirdAlpha dv2(p)
 * (doAligned: the two-pass aligned blend), 3 volumes: the chained blend with
 * fourthAlpha; then the lookup with alpha 1 over the bnd-10 interior. */
int flof_advect_time_slice_vols(flof_grid4d *const *vols, int nvols,
                                float *dst_dev, int dx, int dy, int dz,
                                flof_grid4d *phi, float time, float blendAlpha,
                                float loadTimeScale, const float defoOffset[4],
                                const float defoScale[4],
                                const float defoFactor[4],
                                const float overrideSize[4],
                                float overrideTimeOff, int zeroVel, int bordSkip,
                                float defoAniFac, int doAligned,
                                float partialLoadFac, float thirdAlpha,
                                float fourthAlpha);
/* Batched form: warps nframes output frames (times time0 + f*dtime) in one
 * launch; frame f is written at dst_dev + f*dx*dy*dz. */
int flof_advect_time_slices_opt(flof_grid4d *defo, float *dst_dev, int dx,
                                int dy, int dz, int nframes, float time0,
                                float dtime, flof_grid4d *phi,
                                float blendAlpha, float loadTimeScale,
                                const float defoOffset[4],
                                const float defoScale[4],
                                const float defoFactor[4], int bordSkip,
                                float defoAniFac);

/* ---- input preprocessing (optflow4d.cpp:1254-1266,1361-1406,1598-1616) -- */
/* repeatFrame4d(phi, srct, range=0, bnd=0). */
int flof_repeat_frame4d(flof_grid4d *phi, float srct, float range, int bnd);
/* extrap4dLsSimple(phi, distance=4, inside=false). */
int flof_extrap4d_ls_simple(flof_grid4d *phi, int distance, int inside);
/* shiftForwGrid4d(phi, overrideGoodRegion=0). */
int flof_shift_forw_grid4d(flof_grid4d *phi, int overrideGoodRegion);

/* ---- .uni files (fileio.cpp:38-44,572-599,655-731,834-999) ------------- */
/* getUniFileSize: X, Y, Z of an MNT2 (3D) or M4T2 (4D) file, T (1 for 3D). */
int flof_uni_file_size(const char *name, int dims[4]);
/* writeGrid4dUni from host memory (M4T2, gzip level 1). */
int flof_uni_write4d(const char *name, const void *host, int X, int Y, int Z,
                     int T, int elem);
/* readGridUni / readGrid4dUni into host memory; dims and element type must
 * match the file (T = 1 for a 3D file). */
int flof_uni_read(const char *name, void *host, int X, int Y, int Z, int T,
                  int elem);
/* Grid4d::save / Grid4d::load (grid4d.cpp:122-138). */
int flof_grid4d_save(flof_grid4d *g, const char *name);
int flof_grid4d_load(flof_grid4d *g, const char *name);
/* loadVelRescaled (scenes/ofHelpers.py:91-100): the VEC4 deformation file
 * `fname` (M4T2) is read into a grid of the deformation solver's size
 * defoSize (NULL: the file's size; a file of another size is an error, as in
 * readGrid4dUni, fileio.cpp:936-951), resampled into vel by
 * interpolateGrid4dVec(vel, ., offset=defoOffset, scale=defoScale)
 * (grid4d.cpp:549-557) and multiplied by defoFactor (Grid4d::multConst). */
int flof_load_vel_rescaled(flof_grid4d *vel, const char *fname, const int defoSize[4],
                           const float defoOffset[4], const float defoScale[4],
                           const float defoFactor[4]);

/* ---- file-based frame functions (optflow4d.cpp:1464-1595,1871-2105) ----- */
/* loadPlaceGrid4d(fname, ...): `fname` is the reference's snprintf pattern
 * with one integer conversion; frames are read from MNT2 .uni files. */
int flof_load_place_grid4d_uni(const char *fname, flof_grid4d *phi,
                               const float offset[4], const float scale[4],
                               int fileIdxStart, int fileIdxEnd,
                               int debugSkipLoad, float spread,
                               const float overrideSize[4],
                               float overrideTimeOff, int overrideGoodRegion,
                               float loadTimeScale, int rescaleSdfValues,
                               float sdfIsoOff, float repeatStartFrame);
/* loadAdvectTimeSlice_OptInit(ID, fname1, useDefoVols, doAligned,
 * partialLoadFac) (optflow4d.cpp:1871-1913): loads the M4T2 deformation into
 * device memory; the handle replaces the reference's int-keyed global map. */
int flof_warp_state_open(flof_ctx *ctx, const char *fname, int useDefoVols,
                         int doAligned, float partialLoadFac,
                         flof_warp_state **out);
/* loadAdvectTimeSlice_OptAdd (optflow4d.cpp:1914-1928): another deformation
 * volume (ignored without useDefoVols, as in the reference). */
int flof_warp_state_add(flof_warp_state *s, const char *fname);
/* loadAdvectTimeSlice_OptRun (optflow4d.cpp:1951-2105) with the state's
 * deformation(s); arguments in the reference's order. */
int flof_warp_state_run(flof_warp_state *s, float *dst_dev, int dx, int dy,
                        int dz, flof_grid4d *phi, float time, float blendAlpha,
                        float loadTimeScale, const float defoOffset[4],
                        const float defoScale[4], const float defoFactor[4],
                        const float overrideSize[4], float overrideTimeOff,
                        int zeroVel, float thirdAlpha, int bordSkip,
                        float fourthAlpha, float defoAniFac);
/* loadAdvectTimeSlice_Finish. */
int flof_warp_state_close(flof_warp_state *s);

/* ---- learned warm start (no reference counterpart) ---------------------- */
/* A space-time convolutional network (lift 2->16, four residual blocks of a
 * 3x3x3 space and a 3-tap time convolution with 16 channels, head 16->4)
 * that predicts a 4D flow from the two SDFs; BF16 weights on the matrix
 * cores with FP32 accumulation, FP32 master weights trained with Adam.
 * Parameters are a flat float array of flof_warmnet_num_params() values. */
int flof_warmnet_num_params(void);
int flof_warmnet_create(flof_ctx *ctx, uint64_t seed, flof_warmnet **out);
int flof_warmnet_destroy(flof_warmnet *net);
int flof_warmnet_get_params(flof_warmnet *net, float *host, int n);
int flof_warmnet_set_params(flof_warmnet *net, const float *host, int n);
/* Gradient of the last flof_warmnet_loss_grad call. */
int flof_warmnet_get_grads(flof_warmnet *net, float *host, int n);
/* flow (VEC4, cells of the input grid) = net(i0, i1).  The training entries
 * (forward, loss_grad) keep every activation for the backward pass and take
 * grids below 2^31/16 cells; the solver's own inference (the _net solve)
 * takes larger ones. */
int flof_warmnet_forward(flof_warmnet *net, flof_grid4d *i0, flof_grid4d *i1,
                         flof_grid4d *flow);
/* Mean squared error against target (VEC4) and its gradient (kept in the
 * net for flof_warmnet_adam / flof_warmnet_get_grads). */
int flof_warmnet_loss_grad(flof_warmnet *net, flof_grid4d *i0, flof_grid4d *i1,
                           flof_grid4d *target, double *loss);
/* One Adam step with the last gradient. */
int flof_warmnet_adam(flof_warmnet *net, float lr, float beta1, float beta2,
                      float eps);
/* Warm start for opticalFlowMultiscale4d: i0, i1 are restricted along the
 * solve's pyramid to the first level with X <= netX, the net runs there and
 * its flow is prolonged (x2 per level) into warm (VEC4, size of i0). */
int flof_warmnet_warm_start(flof_warmnet *net, flof_grid4d *i0,
                            flof_grid4d *i1, int netX, flof_grid4d *warm);

/* ---- synthetic inputs (no reference counterpart) ----------------------- */
/* Seeded falling-drop SDF (replaces the FLIP input simulation of
 * scenes/dataGen2Drop.py): a pool plane at y = 0.2*Y united (min) with
 * `nspheres` spheres drawn from the splitmix64 stream `seed` -- centre x in
 * [0.3X,0.7X], start height in [0.6Y,0.8Y], centre z in [0.3Z,0.7Z], radius
 * rmin*X (or uniform in [rmin,rmax]*X) -- that fall by 0.2*Y*s^2 with
 * s = t/(T-1); values in cells, clamped to +-clampCells, computed in FP64 and
 * stored as FP32. */
int flof_drop_sdf4d(flof_grid4d *phi, uint64_t seed, int nspheres, double rmin,
                    double rmax, double clampCells);
/* The same drops as T full-resolution 3D frames at `scale` times the spatial
 * resolution of an X*Y*Z*T grid, in fine-cell units (clamp clampCells*scale).
 * Frame t is written at frames_dev + t*(scale^3*X*Y*Z). */
int flof_drop_frames3d(flof_ctx *ctx, float *frames_dev, int X, int Y, int Z,
                       int T, int scale, uint64_t seed, int nspheres,
                       double rmin, double rmax, double clampCells);
/* sizeof of the ABI structs (0: flof_of_params, 1: flof_solve_info), for
 * binding checks. */
size_t flof_abi_struct_size(int which);

/* ---- device memory helpers for 3D frames ------------------------------- */
int flof_dev_alloc(flof_ctx *ctx, size_t bytes, void **out);
int flof_dev_free(flof_ctx *ctx, void *ptr);
int flof_dev_upload(flof_ctx *ctx, void *dst_dev, const void *src,
                    size_t bytes);
int flof_dev_download(flof_ctx *ctx, void *dst, const void *src_dev,
                      size_t bytes);

#ifdef __cplusplus
}
#endif

#endif /* FLOF_H */
