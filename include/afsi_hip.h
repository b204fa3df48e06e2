/*
 * afsi_hip: MI355X (gfx950) kernels for the MAC/IB per-timestep hot path.
 *
 * C ABI: raw device pointers, sizes and a hipStream_t passed as void*.
 * Every entry point returns 0 on success or a nonzero status; the message
 * is available from afsi_last_error() on the calling thread. Kernels run on
 * the given stream and never synchronise it. Device-side validation (point
 * support, index range, non-finite values) ORs bits into an optional
 * caller-owned device int (`flags`, may be NULL):
 *   1 support violated, 2 index out of range, 4 non-finite, 8 capacity.
 *
 * Reference citations are relative to loveIroha/AFSI_GPU src/afsi_torch.
 */
#ifndef AFSI_HIP_H
#define AFSI_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Uniform 3D MAC grid. `center_origin` = origin + 0.5*h evaluated by the
 * caller in FP64 exactly as mac/grid.py:38-40 face_origin does, so integer
 * stencil bases are bit-identical to the reference's. */
typedef struct afsi_grid3 {
    int64_t n[3];             /* cells per axis */
    double h[3];              /* spacing = length / n (mac/grid.py:27-29) */
    double origin[3];         /* box origin; face lattice along its own axis */
    double center_origin[3];  /* origin + 0.5*h; cell-centred lattice */
    int32_t periodic;         /* 0: closed Neumann box (reference); 1: periodic */
    int32_t reserved;
} afsi_grid3;

const char* afsi_last_error(void);
int afsi_abi_version(void);

/* ---- Stencil preparation -------------------------------------------------
 * Replaces mac/compact_transfer.py:21-26 (_prepare) as used by
 * CompactFETransfer.from_points (:144-150): base (3,P,3) int64 and phi
 * (3,P,3,4) for the three staggered lattices, phi with the reference's two-
 * branch formula (ib.py:57-67). Sets flag 1 for points failing the support
 * predicate of mac/compact_transfer.py:94-96,115-117 (closed box only). */
int afsi_prepare_compact_f64(const double* x, int64_t points, const afsi_grid3* grid,
                             int64_t* base, double* phi, int* flags, void* stream);

/* ---- Compact separable transfer (one component per call) ----------------
 * Replace torch.ops.afsi_ib_cuda.compact_gather / compact_spread
 * (mac/csrc/transfer.cpp:34-53; mac/cuda_ib.py:79-90).
 * field/out are the flat C-order face array of shape (nx,ny,nz) of this
 * component; base (P,3) int64, phi (P,3,4). The gather writes out[p*stride]
 * (one thread per point, separable sums). The spread reads
 * force[p*stride] and writes every entry of `out` (a fresh result) with
 * sum (w*f)/volume over the points whose stencil reaches it: the points are
 * stably sorted by stencil base and each face sums its 4x4 base columns in
 * a fixed order, so the result is deterministic and uses no atomics. The
 * spread needs `workspace` of afsi_compact_spread_workspace_bytes(). A
 * stencil reaching outside the face array sets flag 2 (the reference's
 * device assert): the gather keeps its in-range taps, the spread drops the
 * point. */
int afsi_compact_gather_f64(const double* field, const int64_t* base, const double* phi,
                            int64_t points, int64_t nx, int64_t ny, int64_t nz,
                            double* out, int64_t out_stride, int* flags, void* stream);
int afsi_compact_gather_f32(const float* field, const int64_t* base, const float* phi,
                            int64_t points, int64_t nx, int64_t ny, int64_t nz,
                            float* out, int64_t out_stride, int* flags, void* stream);
size_t afsi_compact_spread_workspace_bytes(int64_t points, int64_t nx, int64_t ny, int64_t nz);
int afsi_compact_spread_f64(const double* force, int64_t force_stride, const int64_t* base,
                            const double* phi, int64_t points, int64_t nx, int64_t ny,
                            int64_t nz, double volume, double* out, void* workspace,
                            size_t workspace_bytes, int* flags, void* stream);
int afsi_compact_spread_f32(const float* force, int64_t force_stride, const int64_t* base,
                            const float* phi, int64_t points, int64_t nx, int64_t ny,
                            int64_t nz, double volume, float* out, void* workspace,
                            size_t workspace_bytes, int* flags, void* stream);

/* ---- Generic indexed transfer --------------------------------------------
 * Replace torch.ops.afsi_ib_cuda.indexed_gather / indexed_spread
 * (mac/csrc/transfer.cpp:19-33,48-53; mac/cuda_ib.py:69-76): field (N,C),
 * ids (P,K) int64, signed weights (P,K). The gather sums a point's links in
 * order (one thread per point and component). The spread writes out
 * (size,C) = sum (f*w)/volume over the links into each node: links are
 * stably sorted by node and each node sums its links in link order
 * (deterministic, no atomics); it needs `workspace` of
 * afsi_indexed_spread_workspace_bytes(). Out-of-range ids set flag 2 and
 * are skipped. */
int afsi_indexed_gather_f64(const double* field, int64_t nodes, int64_t comps,
                            const int64_t* ids, const double* weights, int64_t points,
                            int64_t neighbors, double* out, int* flags, void* stream);
int afsi_indexed_gather_f32(const float* field, int64_t nodes, int64_t comps,
                            const int64_t* ids, const float* weights, int64_t points,
                            int64_t neighbors, float* out, int* flags, void* stream);
size_t afsi_indexed_spread_workspace_bytes(int64_t points, int64_t neighbors, int64_t size);
int afsi_indexed_spread_f64(const double* field, int64_t points, int64_t comps,
                            const int64_t* ids, const double* weights, int64_t neighbors,
                            int64_t size, double volume, double* out, void* workspace,
                            size_t workspace_bytes, int* flags, void* stream);
int afsi_indexed_spread_f32(const float* field, int64_t points, int64_t comps,
                            const int64_t* ids, const float* weights, int64_t neighbors,
                            int64_t size, double volume, float* out, void* workspace,
                            size_t workspace_bytes, int* flags, void* stream);

/* ---- Point binning (build-own; no reference counterpart) ------------------
 * Integer cell of each point = face-lattice base + 1 on every axis
 * (the base of mac/compact_transfer.py:21-23), wrapped on a periodic grid.
 * Keys use a tile-major cell numbering (tile tx*ty*tz, both levels x-major).
 * Output: keys_sorted (P) uint32, perm (P) int32 = stable argsort of the
 * keys, cell_ptr (key_space+1) int32 CSR offsets. Stable LSD radix sort. */
int64_t afsi_bin_key_space(const afsi_grid3* grid, const int32_t tile[3]);
size_t afsi_bin_workspace_bytes(int64_t points);
int afsi_bin_points(const double* x, int64_t points, const afsi_grid3* grid,
                    const int32_t tile[3], uint32_t* keys_sorted, int32_t* perm,
                    int32_t* cell_ptr, void* workspace, size_t workspace_bytes,
                    int* flags, void* stream);

/* ---- Tiled transfer over the binned CSR (build-own fast path) -------------
 * Same mathematics as compact_spread/compact_gather for all three
 * components at once, weights computed from positions in registers.
 * Spread: one work item per (tile, chunk of points); per-tile LDS
 * accumulation of all three components in 64-bit fixed point (a power-of-two
 * scale per work item from its largest force, integer LDS atomics; FP64 LDS
 * atomics for items with nonfinite or extreme forces), then one FP64 global
 * atomic per touched face. u,v,w are overwritten (zero-filled first).
 * Interp: the same work items; each stages its tile's face box of all three
 * components in LDS and gathers its points' 192 values from there; out (P,3)
 * in the caller's point order. With perm = NULL (unbinned) one thread per
 * point gathers from HBM and tile/cell_ptr/workspace are unused. Cubic
 * tiles of edge 8 or 4 use compile-time box extents. */
size_t afsi_spread_workspace_bytes(const afsi_grid3* grid, const int32_t tile[3]);
/* Host only (no device work): the smallest and largest face offset, per
 * component, that any tile's LDS box entry maps to in the tiled kernels;
 * lets a caller check that every staged read and flushed write stays inside
 * the face arrays for its grid and tile (build-own; no reference counterpart). */
int afsi_tile_box_bounds(const afsi_grid3* grid, const int32_t tile[3], int64_t lo[3], int64_t hi[3]);
int afsi_spread_tiled_f64(const double* x, const double* force, int64_t points,
                          const afsi_grid3* grid, const int32_t tile[3],
                          const int32_t* perm, const int32_t* cell_ptr,
                          double* u, double* v, double* w, void* workspace,
                          size_t workspace_bytes, int* flags, void* stream);
/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops): afsi_spread_tiled_f64
 * with no workspace argument. Its one scratch buffer, the per-tile work-list
 * prefix, is a fixed device array of 2^20 ints (4 MiB) in the library, one
 * per GPU, held until the process ends; no hipMalloc/hipFree is called.
 * Grids of more than 2^20 - 1 tiles are rejected with an error. Calls that
 * overlap on one GPU share the array and race. */
int afsi_spread_tiled_noalloc_f64(const double* x, const double* force, int64_t points,
                                  const afsi_grid3* grid, const int32_t tile[3], const int32_t* perm,
                                  const int32_t* cell_ptr, double* u, double* v, double* w, int* flags,
                                  void* stream);
/* The same spread as a per-face-tile gather (build-own): the workgroup that
 * owns a tile produces all its faces from the points binned in the tile and
 * its two-cell halo (same-cell runs accumulated in registers per x-slot,
 * folded into an LDS tile box) and writes each face once with a plain
 * store; no global atomics, no zero fill. Tiles whose halo holds more
 * points than one work item takes are split and their chunk boxes added in
 * chunk order by a second launch. Slower than afsi_spread_tiled_f64 on the
 * clustered-shells workload (1.62 vs 0.28 ms at 128^3 with 1.5M points).
 * The workspace depends on the point count. */
size_t afsi_spread_gather_workspace_bytes(const afsi_grid3* grid, const int32_t tile[3], int64_t points);
int afsi_spread_gather_f64(const double* x, const double* force, int64_t points, const afsi_grid3* grid,
                           const int32_t tile[3], const int32_t* perm, const int32_t* cell_ptr, double* u,
                           double* v, double* w, void* workspace, size_t workspace_bytes, int* flags,
                           void* stream);
int afsi_interp_tiled_f64(const double* x, int64_t points, const afsi_grid3* grid,
                          const int32_t tile[3], const int32_t* perm, const int32_t* cell_ptr,
                          const double* u, const double* v, const double* w, double* out,
                          void* workspace, size_t workspace_bytes, int* flags, void* stream);

/* ---- Batched transfers (build-own) ------------------------------------------
 * Many 3D instances binned, spread and interpolated with one launch per
 * phase: the instances' tile-major key spaces are concatenated (instance
 * i's keys offset by the key spaces before it) and sorted together, and
 * spread/interpolation walk one work list of (instance, tile, chunk of
 * points) with the per-instance LDS boxes of afsi_spread_tiled_f64 /
 * afsi_interp_tiled_f64. perm holds global point ids (instance i's points
 * are numbered after those of instances 0..i-1). Faces are overwritten by
 * the spread; out receives (P_i,3) velocities in each instance's order. */
typedef struct afsi_tb_item {
    const double* x;      /* (points,3) positions */
    const double* force;  /* (points,3) forces */
    double* face[3];      /* face arrays: spread output, interpolation input */
    double* out;          /* (points,3) interpolated velocities */
    int64_t points;
    afsi_grid3 grid;
} afsi_tb_item;
typedef struct afsi_tb_batch {
    int32_t count;
    int32_t tile[3];
    const afsi_tb_item* items;  /* host array */
    uint32_t* keys_sorted;      /* device, total points */
    int32_t* perm;              /* device, total points */
    int32_t* cell_ptr;          /* device, total key space + 1 */
    void* workspace;            /* device */
    size_t workspace_bytes;
    int* flags;                 /* device: 1 support, 2 index, 4 non-finite */
} afsi_tb_batch;
int afsi_tb_sizes(const afsi_tb_batch* batch, int64_t* total_points, int64_t* key_space,
                  size_t* workspace_bytes);
/* Builds the descriptors in the workspace (synchronises the stream). */
int afsi_tb_prepare(const afsi_tb_batch* batch, void* stream);
int afsi_tb_bin(const afsi_tb_batch* batch, void* stream);
int afsi_tb_spread(const afsi_tb_batch* batch, void* stream);
int afsi_tb_interp(const afsi_tb_batch* batch, void* stream);

/* ---- 3D pressure multigrid kernels ---------------------------------------
 * Replace the MG kernel protocol of mac/mg_workspace.py:13-30 /
 * mac/_triton_mg.py:121-151 (smooth, residual, restrict, residual_restrict,
 * prolong_add). n = cells of the level, h2 = h*h per axis (operator
 * (p_nb-p)/h2 in the order of mac/_triton_mg.py:26-32), inv_h2 = 1/h**2
 * (diagonal of mac/multigrid.py:38-41). periodic=1 wraps (build-own). */
int afsi_mg_residual_3d_f64(const double* p, const double* rhs, double* out,
                            const int64_t n[3], const double h2[3], int periodic, void* stream);
int afsi_mg_residual_3d_f32(const float* p, const float* rhs, float* out,
                            const int64_t n[3], const double h2[3], int periodic, void* stream);
/* Weighted Jacobi out = p + omega*(rhs-A p)/diag (mac/multigrid.py:107-110). */
int afsi_mg_jacobi_3d_f64(const double* p, const double* rhs, double* out, const int64_t n[3],
                          const double h2[3], const double inv_h2[3], double omega,
                          int periodic, void* stream);
int afsi_mg_jacobi_3d_f32(const float* p, const float* rhs, float* out, const int64_t n[3],
                          const double h2[3], const double inv_h2[3], double omega,
                          int periodic, void* stream);
/* One colour of red-black Gauss-Seidel in place; colour = (i+j+k)%2
 * (build-own smoother; no reference counterpart). */
int afsi_mg_rbgs_3d_f64(double* p, const double* rhs, const int64_t n[3], const double h2[3],
                        const double inv_h2[3], int color, int periodic, void* stream);
int afsi_mg_rbgs_3d_f32(float* p, const float* rhs, const int64_t n[3], const double h2[3],
                        const double inv_h2[3], int color, int periodic, void* stream);
/* 8-child average (mac/_triton_mg.py:52-68); nc = coarse cells. */
int afsi_mg_restrict_3d_f64(const double* fine, double* coarse, const int64_t nc[3], void* stream);
int afsi_mg_restrict_3d_f32(const float* fine, float* coarse, const int64_t nc[3], void* stream);
/* avg-pool of rhs-A p (mac/_triton_mg.py:71-86); n = fine cells. */
int afsi_mg_residual_restrict_3d_f64(const double* p, const double* rhs, double* coarse,
                                     const int64_t n[3], const double h2[3], int periodic,
                                     void* stream);
int afsi_mg_residual_restrict_3d_f32(const float* p, const float* rhs, float* coarse,
                                     const int64_t n[3], const double h2[3], int periodic,
                                     void* stream);
/* fine += trilinear(coarse) (mac/_triton_mg.py:89-118); n = fine cells. */
int afsi_mg_prolong_add_3d_f64(const double* coarse, double* fine, const int64_t n[3],
                               int periodic, void* stream);
int afsi_mg_prolong_add_3d_f32(const float* coarse, float* fine, const int64_t n[3],
                               int periodic, void* stream);
/* y = A x for a dense row-major (n x n) FP64 matrix: the coarse solve
 * coarse_inverse @ rhs of mac/multigrid.py:114 and mac/mg_workspace.py:86-90. */
int afsi_dense_matvec_f64(const double* A, const double* x, double* y, int64_t n, void* stream);
int afsi_dense_matvec_f32(const float* A, const float* x, float* y, int64_t n, void* stream);

/* ---- Reductions and vector helpers ----------------------------------------
 * Deterministic two-stage reduction: result[0]=sum, [1]=sum of squares,
 * [2]=sum |x|, [3]=count of non-finite entries, all FP64. `partials` is
 * caller scratch of afsi_reduce_partials(count) * 4 doubles. */
int64_t afsi_reduce_partials(int64_t count);
int afsi_reduce_f64(const double* x, int64_t count, double* partials, double* result,
                    void* stream);
int afsi_reduce_f32(const float* x, int64_t count, double* partials, double* result,
                    void* stream);
/* x -= sum[0]/count (mean removal with a device-resident sum). */
int afsi_sub_mean_f64(double* x, int64_t count, const double* sum, void* stream);
int afsi_sub_mean_f32(float* x, int64_t count, const double* sum, void* stream);
int afsi_fill_f64(double* x, int64_t count, double value, void* stream);
int afsi_fill_f32(float* x, int64_t count, double value, void* stream);

/* ---- MAC projection pieces (mac/flow.py:66-71,101-120; mac/grid.py:70-83)
 * out = scale * D(u,v,w), D = sum_c diff(u_c)/h_c; scale = -rho/dt. */
int afsi_divergence_3d_f64(const double* u, const double* v, const double* w, double* out,
                           const int64_t n[3], const double h[3], double scale, int periodic,
                           void* stream);
/* u_c -= coef * G_c p (coef = dt/rho); Neumann boundary-normal faces keep
 * zero gradient (mac/grid.py:74-83). */
int afsi_grad_correct_3d_f64(double* u, double* v, double* w, const double* p,
                             const int64_t n[3], const double h[3], double coef, int periodic,
                             void* stream);
/* y += a * x (the predictor u* = u + (dt/rho) f of the IB step). */
int afsi_axpy_f64(double* y, const double* x, int64_t count, double a, void* stream);
/* The IB step's projection with the predictor u* = u + a f folded in
 * (mac/coupling.py:80-83 then mac/flow.py:117-119), leaving u untouched
 * until the pressure is known:
 *   afsi_project_rhs_3d_f64:     out = scale * D(u + a f) (u, f read only);
 *   afsi_project_correct_3d_f64: u_c = (u_c + a f_c) - coef * G_c p.
 * Each face value u + a f is formed exactly as afsi_axpy_f64 forms it. */
int afsi_project_rhs_3d_f64(const double* u, const double* v, const double* w, const double* fu,
                            const double* fv, const double* fw, double a, double* out,
                            const int64_t n[3], const double h[3], double scale, int periodic,
                            void* stream);
int afsi_project_correct_3d_f64(double* u, double* v, double* w, const double* fu,
                                const double* fv, const double* fw, double a, const double* p,
                                const int64_t n[3], const double h[3], double coef, int periodic,
                                void* stream);

/* ---- Native V-cycle driver -------------------------------------------------
 * Replaces MGWorkspace.cycle/_block/residual_norm (mac/mg_workspace.py:
 * 70-133): issues every launch of `cycles` V-cycles on one stream from C++
 * (capturable as one hipGraph). Per level: nu sweeps (Jacobi ping-pong with
 * tmp, or in-place red-black GS), mean removal, fused residual-restrict,
 * coarse mean removal, zero coarse iterate, recurse, prolong-add, nu sweeps,
 * mean removal; the coarsest level is coarse_inverse @ rhs minus its mean
 * (mac/multigrid.py:112-123). Buffers belong to the caller; `T` of p/rhs/
 * tmp/coarse_inverse is the function's precision. reduce_partials holds
 * afsi_reduce_partials(level-0 cells)*4 doubles, reduce_result 4 doubles. */
#define AFSI_MG_MAX_LEVELS 16
typedef struct afsi_mg3d_plan {
    int32_t levels;
    int32_t periodic;
    int32_t smoother;        /* 0: weighted Jacobi (omega); 1: red-black GS */
    int32_t nu;              /* sweeps before and after the coarse correction */
    double omega;
    int64_t n[AFSI_MG_MAX_LEVELS][3];
    double h2[AFSI_MG_MAX_LEVELS][3];
    double inv_h2[AFSI_MG_MAX_LEVELS][3];
    void* p[AFSI_MG_MAX_LEVELS];
    void* rhs[AFSI_MG_MAX_LEVELS];
    void* tmp[AFSI_MG_MAX_LEVELS];
    const void* coarse_inverse;
    int64_t coarse_cells;
    double* reduce_partials;
    double* reduce_result;
    int64_t partials_capacity; /* doubles in reduce_partials */
    int32_t zero_initial;      /* fused driver: treat p[0] as zero on the first cycle */
    /* red-black, nu = 2: levels with at most fuse_cells cells (0: 262144;
     * negative: none) and z extent 32 or 64 run their pre-smoothing,
     * residual and restriction as one launch and their prolongation and
     * post-smoothing as another (tiles recompute their halo rows). */
    int32_t fuse_cells;
    /* fused red-black sweeps, per level (0 = automatic): tile edge TX = TY
     * (2 or 4) and workgroups per launch (persistent: each workgroup sweeps
     * every G-th tile, overlapping the next tile's loads with compute). */
    int32_t sweep_tile[AFSI_MG_MAX_LEVELS];
    int32_t sweep_blocks[AFSI_MG_MAX_LEVELS];
    /* red-black over-relaxation: each red-black update p <- gs becomes
     * p <- p + relax * (gs - p) (build-own; 0 or 1: plain Gauss-Seidel,
     * the reference's red-black composition). */
    double relax;
} afsi_mg3d_plan;
/* 1 if the fused driver accepts the plan (every level with <= 4096 cells
 * that it solves in LDS must have extents <= 64), else 0. */
int afsi_mg3d_supported(const afsi_mg3d_plan* plan);
/* Doubles of reduce_partials the fused driver needs for level-0 extent n. */
int64_t afsi_mg3d_partials_doubles(const int64_t n[3]);
/* Fused driver (the production path). Per fine level: nu single-launch
 * sweeps (red-black: LDS tile + 2-cell halo, red on tile+1 ring then
 * black; or weighted Jacobi), one residual-restriction, one prolongation;
 * levels selected by fuse_cells instead run two launches (sweeps +
 * residual + restriction; prolongation + sweeps) with identical results.
 * All levels with <= 4096 cells run as one 1024-thread workgroup out of
 * LDS (including the dense coarse solve). No level performs the
 * reference's intermediate mean subtractions (mac/multigrid.py:107-123):
 * the V-cycle commutes with adding a constant, so they change only
 * rounding; p[0] is returned with an arbitrary constant offset that the
 * caller removes once after the last cycle. */
int afsi_mg3d_cycles_f64(const afsi_mg3d_plan* plan, int32_t cycles, void* stream);
int afsi_mg3d_cycles_f32(const afsi_mg3d_plan* plan, int32_t cycles, void* stream);
/* reduce_result = [sum, sum of squares, sum |.|, non-finite count] of
 * rhs[0] - A p[0] (two launches, no grid-wide synchronisation). */
int afsi_mg3d_residual_f64(const afsi_mg3d_plan* plan, void* stream);
int afsi_mg3d_residual_f32(const afsi_mg3d_plan* plan, void* stream);
/* One smoothing sweep of `level` from p[level] into tmp[level] (exactly
 * the kernel the fused driver launches there); for timing. */
int afsi_mg3d_sweep_f64(const afsi_mg3d_plan* plan, int32_t level, void* stream);
int afsi_mg3d_sweep_f32(const afsi_mg3d_plan* plan, int32_t level, void* stream);
/* FP64 iterative refinement around FP32 V-cycles (build-own mixed
 * precision): r = b - A x in FP64 (difference form), r32 = float(r), and
 * result = [sum r, sum r^2, sum |r|, non-finite] (two launches);
 * partials holds partials_capacity doubles. */
int afsi_refine_residual_3d(const double* x, const double* b, float* r32, const int64_t n[3],
                            const double inv_h2[3], int periodic, double* partials,
                            int64_t partials_capacity, double* result, void* stream);
/* One fused refinement step: x_out = x + e and r32 = float(b - A x_out) in
 * FP64 (neighbour values formed from the same x + e expression; identical
 * to afsi_refine_update followed by afsi_refine_residual_3d). x_out must not
 * alias x. result = NULL skips the reduction; otherwise result receives
 * [sum r, sum r^2, sum |r|, non-finite] (partials as above). */
int afsi_refine_step_3d(const double* x, const float* e, const double* b, double* x_out,
                        float* r32, const int64_t n[3], const double inv_h2[3], int periodic,
                        double* partials, int64_t partials_capacity, double* result, void* stream);
/* x += e (FP64 accumulation of an FP32 correction). */
int afsi_refine_update(double* x, const float* e, int64_t count, void* stream);
/* Unfused driver with the reference workspace's launch structure (one
 * launch per colour sweep, separate reductions and mean subtractions,
 * mac/mg_workspace.py:70-133); kept as a cross-check of the fused one.
 * tmp[0] receives the residual in the _residual_ variant. */
int afsi_mg3d_cycles_unfused_f64(const afsi_mg3d_plan* plan, int32_t cycles, void* stream);
int afsi_mg3d_cycles_unfused_f32(const afsi_mg3d_plan* plan, int32_t cycles, void* stream);
int afsi_mg3d_residual_unfused_f64(const afsi_mg3d_plan* plan, void* stream);
int afsi_mg3d_residual_unfused_f32(const afsi_mg3d_plan* plan, void* stream);

/* ==== Batched mixed-precision pressure solve (build-own) ====================
 * Many independent 3D instances of uneven size solved together: one launch
 * per multigrid phase covers every instance (a device work list of
 * (instance, tile/row) entries; each workgroup finds its entry by binary
 * search), the coarse levels run one 1024-thread workgroup per instance,
 * and convergence is tracked on the device (active[i] = 0 once instance i
 * meets max(atol, rtol ||b_i||); inactive instances skip their work).
 * Each item: the FP32 plan of its inner V-cycle (as for
 * afsi_mg3d_cycles_f32, zero initial guess, red-black smoother, every
 * level above the LDS-solved ones with even z extent <= 128), its FP64
 * iterates x[2] (ping-pong) and FP64 right-hand side b. One outer
 * iteration = one batched FP32 V-cycle on r32 = plan.rhs[0], then
 * x_new = x + plan.p[0] and the FP64 residual (afsi_refine_step_3d). */
typedef struct afsi_mgb_item {
    afsi_mg3d_plan plan;
    double* x[2];
    double* b;
} afsi_mgb_item;
typedef struct afsi_mgb_batch {
    int32_t count;
    int32_t inner_cycles;         /* FP32 V-cycles per refinement iteration (0: 1) */
    const afsi_mgb_item* items;   /* host array of count items */
    void* workspace;              /* device, afsi_mgb_workspace_bytes */
    size_t workspace_bytes;
    int32_t* active;              /* device, count */
    double* tol;                  /* device, count */
    double* results;              /* device, 16 per instance: [sum, sum sq, sum |.|, non-finite]
                                     of the residual, of b before and after its mean removal,
                                     and of the iterate whose mean was last removed */
    double* partials;             /* device, afsi_mgb_partials_doubles */
} afsi_mgb_batch;
size_t afsi_mgb_workspace_bytes(const afsi_mgb_batch* batch);
int64_t afsi_mgb_partials_doubles(const afsi_mgb_batch* batch);
/* Builds the per-launch descriptor arrays and copies them to the workspace
 * (synchronises the stream; call once after the buffers are allocated). */
int afsi_mgb_prepare(const afsi_mgb_batch* batch, void* stream);
/* b_i and x_i[0] lose their means (FP64), tol_i = max(atol, rtol ||b_i||),
 * results[16i..16i+3] = residual stats of x_i[0], results[16i+4..16i+7] =
 * stats of b_i before the mean removal, results[16i+12..] = stats of x_i[0]
 * before its mean removal, active_i = residual > tol_i. */
int afsi_mgb_begin(const afsi_mgb_batch* batch, double rtol, double atol, void* stream);
/* The same from x_i[start] (start 0 or 1): a solve that continues from the
 * iterate the previous one ended on, with no copy. */
int afsi_mgb_begin_at(const afsi_mgb_batch* batch, int32_t start, double rtol, double atol, void* stream);
/* `count` outer iterations from x[start] (the iterate alternates between
 * x[0] and x[1]); the last one reduces the residual and updates active.
 * Inactive instances copy their iterate so the alternation stays aligned. */
int afsi_mgb_iterate(const afsi_mgb_batch* batch, int32_t start, int32_t count, void* stream);
/* x_i[cur] loses its mean (FP64). */
int afsi_mgb_finish(const afsi_mgb_batch* batch, int32_t cur, void* stream);

/* ---- elementwise passes of a batch of 3D IB instances ---------------------
 * Build-own (no single reference counterpart): the per-instance projection
 * pieces of mac/flow.py:102-121 with the IB predictor folded in
 * (afsi_project_rhs_3d_f64 / afsi_project_correct_3d_f64, element for
 * element the same expressions) and the point advance x += dt U, each one
 * launch over all instances. All instances share the boundary type. */
typedef struct afsi_ib3b_item {
    double* velocity[3];          /* face arrays, updated in place */
    const double* force[3];       /* spread face force densities */
    double* rhs;                  /* pressure right-hand side (n cells) */
    const double* pressure[2];    /* the two pressure buffers correct() can read */
    double* x;                    /* (points, 3) positions */
    const double* point_velocity; /* (points, 3) */
    int64_t points;
    afsi_grid3 grid;
} afsi_ib3b_item;
typedef struct afsi_ib3b_batch {
    int32_t count;
    int32_t reserved;
    const afsi_ib3b_item* items;  /* host array of count items */
    void* workspace;              /* device, afsi_ib3b_workspace_bytes */
    size_t workspace_bytes;
} afsi_ib3b_batch;
size_t afsi_ib3b_workspace_bytes(const afsi_ib3b_batch* batch);
/* Copies the descriptors to the workspace (synchronises the stream); call
 * again whenever an item's pointers change. */
int afsi_ib3b_prepare(const afsi_ib3b_batch* batch, void* stream);
/* rhs_i = scale * D(u_i + a f_i), u_i unchanged. */
int afsi_ib3b_rhs(const afsi_ib3b_batch* batch, double a, double scale, void* stream);
/* u_i = (u_i + a f_i) - coef * G pressure_i[which] (closed boxes: boundary
 * faces keep u + a f). */
int afsi_ib3b_correct(const afsi_ib3b_batch* batch, int32_t which, double a, double coef, void* stream);
/* x_i += dt * point_velocity_i. */
int afsi_ib3b_advance(const afsi_ib3b_batch* batch, double dt, void* stream);

/* ==== 2D staggered channel (mac2d) ========================================
 * Cells nx x ny, cell (x, y) at x*ny + y; faces u (nx+1, ny), v (nx, ny+1).
 * periodic = 0: the reference channel (mac2d/grid.py:61-71: Neumann inlet
 * and walls, p = 0 half a cell beyond the outlet, non-singular);
 * periodic = 1 (build-own): periodic in x (u has nx faces), walls in y. */
typedef struct afsi_grid2 {
    int64_t n[2];        /* cells per axis */
    double h[2];         /* spacing = length / n (mac2d/grid.py:19-21) */
    double lengths[2];   /* channel lengths (support test of mac2d/transfer.py:88-95) */
    int32_t periodic;
    int32_t reserved;
} afsi_grid2;

/* Replaces TriangleTransfer._prepare (mac2d/transfer.py:61-81) for points
 * (P,2): per component (P,16) int64 ids and FP64 weights, links k = 4i+j,
 * rows outside the walls reflected with negated weights; periodic grids
 * wrap x. Points outside the support region of mac2d/transfer.py:88-95
 * (periodic: y only) set flag 1 and get zero weights. Spread and gather
 * then use afsi_indexed_spread_* / afsi_indexed_gather_* per component
 * (mac2d/transfer.py:108-127). */
int afsi_channel_prepare_f64(const double* x, int64_t points, const afsi_grid2* grid,
                             int64_t* ids_u, double* w_u, int64_t* ids_v, double* w_v, int* flags,
                             void* stream);

/* Multigrid kernels of the channel (mac2d/_triton_mg.py:24-75, bit-exact
 * in FP64): residual out = rhs - A p; Jacobi out = p + (omega*(rhs-Ap))/diag
 * (mac2d/multigrid.py:10-13, omega = 2/3); red-black GS (build-own, (x+y)
 * even first, in place); residual-restriction to the 2:1 coarse level;
 * ghost-based bilinear prolong-add. n = cells of the (fine) level. */
int afsi_mg_residual_2d_f64(const double* p, const double* rhs, double* out, const int64_t n[2],
                            const double h2[2], int periodic, void* stream);
int afsi_mg_residual_2d_f32(const float* p, const float* rhs, float* out, const int64_t n[2],
                            const double h2[2], int periodic, void* stream);
int afsi_mg_jacobi_2d_f64(const double* p, const double* rhs, double* out, const int64_t n[2],
                          const double h2[2], const double inv_h2[2], double omega, int periodic,
                          void* stream);
int afsi_mg_jacobi_2d_f32(const float* p, const float* rhs, float* out, const int64_t n[2],
                          const double h2[2], const double inv_h2[2], double omega, int periodic,
                          void* stream);
int afsi_mg_rbgs_2d_f64(double* p, const double* rhs, const int64_t n[2], const double h2[2],
                        const double inv_h2[2], int periodic, void* stream);
int afsi_mg_rbgs_2d_f32(float* p, const float* rhs, const int64_t n[2], const double h2[2],
                        const double inv_h2[2], int periodic, void* stream);
int afsi_mg_residual_restrict_2d_f64(const double* p, const double* rhs, double* coarse,
                                     const int64_t n[2], const double h2[2], int periodic,
                                     void* stream);
int afsi_mg_residual_restrict_2d_f32(const float* p, const float* rhs, float* coarse,
                                     const int64_t n[2], const double h2[2], int periodic,
                                     void* stream);
int afsi_mg_prolong_add_2d_f64(const double* coarse, double* fine, const int64_t n[2], int periodic,
                               void* stream);
int afsi_mg_prolong_add_2d_f32(const float* coarse, float* fine, const int64_t n[2], int periodic,
                               void* stream);

/* Channel divergence out = scale*(du/hx + dv/hy) (mac2d/grid.py:46-48) and
 * in-place correction u -= coef*G p with the outlet gradient -2p/hx
 * (mac2d/grid.py:51-58; mac2d/flow.py:88-90). */
int afsi_divergence_2d_f64(const double* u, const double* v, double* out, const int64_t n[2],
                           const double h[2], double scale, int periodic, void* stream);
int afsi_grad_correct_2d_f64(double* u, double* v, const double* p, const int64_t n[2],
                             const double h[2], double coef, int periodic, void* stream);

/* Replaces ChannelWorkspace.advance/residual_norm (mac2d/mg_workspace.py:
 * 55-106) and ChannelMultigrid._cycle (mac2d/multigrid.py:65-81): cycles
 * V(nu,nu) issued on one stream (capturable). Reference channel: no mean
 * removal, coarsest = inverse(A) @ rhs. Periodic: means removed after each
 * smoothing stage, from restricted residuals and from the coarsest solution
 * (inverse(A + 11^T/n)), as mac/multigrid.py:107-123 does in 3D.
 * tmp[l] is the Jacobi ping-pong buffer and, at level 0, the residual. */
typedef struct afsi_mg2d_plan {
    int32_t levels;
    int32_t periodic;
    int32_t smoother;        /* 0: weighted Jacobi (omega); 1: red-black GS */
    int32_t nu;
    double omega;
    int64_t n[AFSI_MG_MAX_LEVELS][2];
    double h2[AFSI_MG_MAX_LEVELS][2];
    double inv_h2[AFSI_MG_MAX_LEVELS][2];
    void* p[AFSI_MG_MAX_LEVELS];
    void* rhs[AFSI_MG_MAX_LEVELS];
    void* tmp[AFSI_MG_MAX_LEVELS];
    const void* coarse_inverse;
    int64_t coarse_cells;
    double* reduce_partials;
    double* reduce_result;
    int64_t partials_capacity; /* doubles in reduce_partials */
} afsi_mg2d_plan;
int afsi_mg2d_cycles_f64(const afsi_mg2d_plan* plan, int32_t cycles, void* stream);
int afsi_mg2d_cycles_f32(const afsi_mg2d_plan* plan, int32_t cycles, void* stream);
/* reduce_result = [sum, sum of squares, sum |.|, non-finite count] of
 * rhs[0] - A p[0]; the residual is left in tmp[0]. */
int afsi_mg2d_residual_f64(const afsi_mg2d_plan* plan, void* stream);
int afsi_mg2d_residual_f32(const afsi_mg2d_plan* plan, void* stream);

/* ---- batched channel instances (build-own; no reference counterpart) ------
 * Many channels of uneven size in one launch per phase. Face arrays of the
 * batch are concatenated (u faces of instance 0, 1, ..., then separately v
 * faces); the stencil kernel writes global face ids into (P_total,16) id and
 * weight arrays (rows in instance order), so one afsi_indexed_spread_f64 /
 * afsi_indexed_gather_f64 per component serves the batch. Each instance's
 * FP64 pressure hierarchy is its own afsi_mg2d_plan: divergence writes its
 * rhs[0], cycles advance every active instance (active[i] != 0; NULL: all)
 * by `cycles` V-cycles: levels of more than 4096 cells of a red-black
 * instance as one launch per phase over all instances that have them, the
 * rest in one 1024-thread workgroup per instance; correct reads its p[0]. stats writes
 * out[4i..4i+3] = [sum, sum of squares, sum |.|, non-finite count] of
 * rhs[0] (which 0), p[0] (1) or rhs[0] - A p[0] (2) in a fixed order;
 * which 3 / 4 subtracts out[4i]/cells from rhs[0] / p[0] of the periodic
 * instances (the reference channel is non-singular). prepare uploads
 * the descriptors into the workspace (synchronises the stream). */
typedef struct afsi_chb_item {
    const double* x;        /* (points,2) positions */
    int64_t points;
    afsi_grid2 grid;
    afsi_mg2d_plan plan;    /* FP64 hierarchy of the instance's pressure solve */
} afsi_chb_item;
typedef struct afsi_chb_batch {
    int32_t count;
    int32_t reserved;
    const afsi_chb_item* items;  /* host array */
    void* workspace;             /* device */
    size_t workspace_bytes;
} afsi_chb_batch;
int afsi_chb_sizes(const afsi_chb_batch* batch, int64_t* points, int64_t* u_faces, int64_t* v_faces,
                   int64_t* cells, size_t* workspace_bytes);
int afsi_chb_prepare(const afsi_chb_batch* batch, void* stream);
int afsi_chb_stencil_f64(const afsi_chb_batch* batch, int64_t* ids_u, double* w_u, int64_t* ids_v, double* w_v,
                         int* flags, void* stream);
int afsi_chb_divergence_f64(const afsi_chb_batch* batch, const double* u, const double* v, double scale,
                            void* stream);
int afsi_chb_correct_f64(const afsi_chb_batch* batch, double* u, double* v, double coef, void* stream);
int afsi_chb_cycles_f64(const afsi_chb_batch* batch, const int32_t* active, int32_t cycles, void* stream);
int afsi_chb_stats_f64(const afsi_chb_batch* batch, int which, double* out, void* stream);
/* fu (u_faces) and fv (v_faces) = the batch's face densities: every link of
 * the stencil kernel's (P_total,16) arrays added, force[p,c] * w, with
 * force (P_total,2) already divided by the cell volume (the two
 * afsi_indexed_spread_f64 calls it replaces). Pairs (point, component) are
 * sorted by stencil-base tile, each tile's links are summed in an FP64 LDS
 * box and added to HBM with one atomic per nonzero entry, so the summation
 * order is not fixed. Unsupported points are skipped (the stencil kernel
 * flags them). workspace: afsi_chb_spread_workspace_bytes(batch) bytes. */
size_t afsi_chb_spread_workspace_bytes(const afsi_chb_batch* batch);
int afsi_chb_spread_f64(const afsi_chb_batch* batch, const double* force, double* fu, double* fv, void* workspace,
                        size_t workspace_bytes, int* flags, void* stream);

/* ---- explicit MAC predictor of the closed box ------------------------------
 * u* = zero_normal(u + dt (-N(u) + (mu/rho) L u + f/rho)) with u =
 * zero_normal(u0); replaces MACFlow._predict (mac/flow.py:59-64) with the
 * convection of mac/grid.py:108-134 and velocity_laplacian of
 * mac/grid.py:96-105. Face arrays of a closed n0 x n1 x n2 box (extent
 * n[c] + 1 along c); out of place; bit-identical to the FP64 restatement. */
int afsi_mac_predict_f64(const double* u0, const double* u1, const double* u2, const double* f0,
                         const double* f1, const double* f2, double* s0, double* s1, double* s2,
                         const int64_t n[3], const double h[3], double dt, double rho, double mu, void* stream);
/* out[0..2] = max |u_c|, out[3] = count of non-finite entries of u and f:
 * the inputs of MACFlow.step's finiteness check and transport screen
 * (mac/flow.py:137-146). */
size_t afsi_mac_speeds_workspace_bytes(void);
int afsi_mac_speeds_f64(const double* u0, const double* u1, const double* u2, const double* f0,
                        const double* f1, const double* f2, const int64_t n[3], double* out,
                        void* workspace, size_t workspace_bytes, void* stream);

/* ---- FE interaction points and the consistent mass ------------------------
 * The quadrature IB equations of mac/transfer.py:1-5: M F = b,
 * f = V^-1 K^T W B F, M U = B^T W K u. K is the tiled spread/interpolation
 * above, applied at the interaction points B x. All arrays FP64 row-major,
 * nodal and point fields (count, 3). */
typedef struct afsi_fe_geometry {
    const int64_t* cells;      /* (elements, nodes_per_cell) node ids */
    const double* values;      /* (quad, nodes_per_cell): N_qa */
    const double* weights;     /* (elements, quad): physical weights W_eq */
    int64_t elements;
    int32_t nodes_per_cell;    /* <= 32 */
    int32_t quad;              /* <= 64 */
    int64_t node_count;
    const int32_t* node_ptr;   /* (node_count + 1): incidences of each node */
    const int32_t* node_items; /* e * nodes_per_cell + a, ascending per node */
} afsi_fe_geometry;
/* B: out[e*quad + q] = sum_a N_qa nodal[cells[e,a]]; replaces
 * FETransfer.evaluate / interaction_points (mac/transfer.py:57-58, 69-71). */
int afsi_fe_evaluate_f64(const afsi_fe_geometry* g, const double* nodal, double* out, void* stream);
/* W B: the same times W_eq; replaces FETransfer.weighted_force
 * (mac/transfer.py:66-67). */
int afsi_fe_weighted_f64(const afsi_fe_geometry* g, const double* nodal, double* out, void* stream);
/* B^T W: out[n] = sum over the incidences (e, a) of n of
 * sum_q N_qa W_eq U[e*quad + q] (deterministic order); replaces
 * FETransfer.assemble_velocity (mac/transfer.py:126-129). */
int afsi_fe_assemble_f64(const afsi_fe_geometry* g, const double* point_values, double* out, void* stream);

typedef struct afsi_csr {
    const int32_t* row_ptr;    /* rows + 1 */
    const int32_t* cols;
    const double* vals;
    int64_t rows;
    int64_t nnz;
} afsi_csr;
/* y = A x for three columns; replaces FETransfer.mass_action
 * (torch.sparse.mm, mac/transfer.py:60-61) and MassSolver.action
 * (mac/mass_solver.py:53-54). */
int afsi_csr_spmv3_f64(const afsi_csr* a, const double* x, double* y, void* stream);

typedef struct afsi_pcg_options {
    double rtol, atol;
    int32_t max_iterations, recompute_every, check_every, reserved;
} afsi_pcg_options;
typedef struct afsi_pcg_info {
    int32_t status;            /* 0 converged, 1 breakdown, 2 not converged,
                                  3 nonfinite true residual */
    int32_t iterations;
    double residual_norm;      /* true residual norm at return */
    double rhs_norm;
    double tolerance;          /* max(atol, rtol * ||rhs||) */
} afsi_pcg_info;
size_t afsi_mass_pcg_workspace_bytes(int64_t rows);
/* Jacobi PCG on A x = rhs (three columns) with the recurrence, check and
 * restart schedule and true-residual acceptance of pcg()
 * (fluid/solvers.py:39-122) and MassSolver.solve (mac/mass_solver.py:56-103);
 * replaces FETransfer.solve_mass (mac/transfer.py:63-64). initial may be
 * null (zero start) or equal x. The device state is read back only at the
 * schedule's check iterations. */
int afsi_mass_pcg_f64(const afsi_csr* a, const double* diagonal, const double* rhs, const double* initial,
                      double* x, const afsi_pcg_options* opt, void* workspace, size_t workspace_bytes,
                      afsi_pcg_info* info, void* stream);

/* ---- P2 fused cell-block transfer ----------------------------------------
 * One quadrature group of 10-node (P2) tetrahedra: every element carries the
 * same reference rule N (quad, 10) and its own physical weights W. */
typedef struct afsi_p2_group {
    const int64_t* cells;      /* (elements, 10) node ids */
    const void* values;        /* (quad, 10): N_qa, the call's element type */
    const void* weights;       /* (elements, quad): W_eq (unused by prepare) */
    int64_t elements;
    int64_t quad;
    int64_t node_count;
} afsi_p2_group;
/* Shared face/centre Peskin tables of mac/shared_stencil.py:8-12:
 * base (2, total, 3) int64, phi (2, total, 3, 4); lattice 0 is faces, 1 is
 * centres; component c reads lattice 0 along axis c and lattice 1 elsewhere.
 * A group occupies rows offset .. offset + elements*quad. */
typedef struct afsi_p2_tables {
    int64_t* base;
    void* phi;                 /* the call's element type */
    int64_t total;
    int64_t offset;
} afsi_p2_tables;
/* Replaces torch.ops.afsi_ib_cuda.p2_prepare_ (mac/csrc/p2_transfer.cpp:50-59,
 * 67; kernel p2_transfer.cu:12-44): interaction points x_q = sum_a N_qa
 * x[cells[e,a]] and their shared tables. ORs 1 into *invalid (device int)
 * when a point is non-finite or 2 <= (x - origin)/h < n - 2 fails on some
 * axis (mac/compact_transfer.py:94-96), 2 for a node id out of range.
 * Closed boxes only (grid->periodic = 0). */
int afsi_p2_prepare_f64(const double* x, const afsi_p2_group* g, const afsi_grid3* grid, afsi_p2_tables* t,
                        int* invalid, void* stream);
/* The same in FP32 (positions, N and tables float; p2_transfer.cu
 * dispatches both types). */
int afsi_p2_prepare_f32(const float* x, const afsi_p2_group* g, const afsi_grid3* grid, afsi_p2_tables* t,
                        int* invalid, void* stream);
/* Replaces torch.ops.afsi_ib_cuda.p2_spread_ (p2_transfer.cpp:66-72,
 * 69; kernel p2_transfer.cu:97-188 with Spread=true): accumulates
 * f_c += (phi_x phi_y phi_z) (W_eq sum_a N_qa value[cells[e,a]]_c) / volume
 * into the face arrays u0, u1, u2 (C-order (nx+1,ny,nz), (nx,ny+1,nz),
 * (nx,ny,nz+1)), which the caller zeroes first. ORs 1 (2) into *flags and
 * skips an element whose stencil box leaves the face arrays (whose node ids
 * are out of range). */
int afsi_p2_spread_f64(const double* value, const afsi_p2_group* g, const afsi_p2_tables* t,
                       const afsi_grid3* grid, double volume, double* u0, double* u1, double* u2, int* flags,
                       void* stream);
/* FP32 fields and tables; products and the LDS patch are accumulated in FP64. */
int afsi_p2_spread_f32(const float* value, const afsi_p2_group* g, const afsi_p2_tables* t,
                       const afsi_grid3* grid, double volume, float* u0, float* u1, float* u2, int* flags,
                       void* stream);
/* Replaces torch.ops.afsi_ib_cuda.p2_gather_assemble_ (p2_transfer.cpp:61-65,
 * 68; kernel p2_transfer.cu:97-188 with Spread=false): accumulates
 * rhs[cells[e,a]]_c += sum_q (N_qa W_eq) sum_taps u_c w into rhs
 * (node_count, 3). Same flags as the spread. */
int afsi_p2_gather_assemble_f64(const double* u0, const double* u1, const double* u2, const afsi_p2_group* g,
                                const afsi_p2_tables* t, const afsi_grid3* grid, double* rhs, int* flags,
                                void* stream);
/* FP32 fields and tables; sums in FP64, rhs accumulated in FP32. */
int afsi_p2_gather_assemble_f32(const float* u0, const float* u1, const float* u2, const afsi_p2_group* g,
                                const afsi_p2_tables* t, const afsi_grid3* grid, float* rhs, int* flags,
                                void* stream);

/* ---- CSR-assembled quadrature operator (P1 meshes) -----------------------
 * Bounded open-addressing table of (key, value) pairs, keys int64 with -1 =
 * empty, capacity a power of two; the caller fills keys with -1 and values
 * and *flag with 0 (mac/hash_transfer_assembly.py:26-50). Buckets are the
 * SplitMix64 finaliser of the key (hash_transfer_assembly.py:17-23), probed
 * linearly up to 128 slots; a full probe or a negative key ORs 1 into
 * *flag, an invalid cell or stencil box ORs 2. */
typedef struct afsi_hash_table {
    int64_t* keys;
    double* values;
    int64_t capacity;
    int* flag;
} afsi_hash_table;
/* Replaces torch.ops.afsi_ib_cuda.ib_accumulate (mac/csrc/ib.cu:110-115,
 * 154-166): values[slot(keys[i])] += values[i], zeros skipped. */
int afsi_hash_accumulate_f64(const int64_t* keys, const double* values, int64_t count, const afsi_hash_table* t,
                             void* stream);
/* Replaces torch.ops.afsi_ib_cuda.ib_contract / hash_contract_
 * (mac/csrc/ib.cu:56-108, 118-152; mac/csrc/ib.cpp:58-65) for one
 * component: adds B_c[node, face] = sum over the cells' interaction points
 * of W_eq N_qa (phi_x phi_y phi_z)(face - x_eq) under the key
 * node * faces_c + face. base (3,P,3) int64 and phi (3,P,3,4) are compact
 * tables of P = lattice_points cell-major interaction points
 * (afsi_prepare_compact_f64); the cells' points start at point `offset`
 * (the reference's per-group offset). Cells of at most 4 nodes; closed
 * boxes only. */
int afsi_p1_contract_f64(const afsi_fe_geometry* g, const int64_t* base, const double* phi, int64_t lattice_points,
                         int64_t offset, int component, const afsi_grid3* grid, const afsi_hash_table* t,
                         void* stream);
/* y[r * y_stride] = scale * sum_k A[r, cols[k]] x[cols[k] * x_stride]: the
 * CSR products of the assembled spread (B^T c / V) and gather (B u) of
 * mac/assembled_transfer.py:199-211 (torch.sparse.mm). */
int afsi_csr_spmv_f64(const afsi_csr* a, const double* x, int64_t x_stride, double* y, int64_t y_stride,
                      double scale, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* AFSI_HIP_H */
