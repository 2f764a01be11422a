/*
 * alamo_elastic.h -- C-ABI of the MI355X nodal elastic multigrid library
 * (libalamo_elastic.so).  It replaces, for one GPU, the call path
 *
 *   Solver::Nonlocal::Newton<T>::solve            src/Solver/Nonlocal/Newton.H:279
 *    -> Solver::Nonlocal::Linear::solve           src/Solver/Nonlocal/Linear.H:131
 *       -> amrex::MLMG -> Operator::Elastic<SYM>  src/Operator/Elastic.H:21-190
 *
 * (paths relative to the reference root).  Conventions:
 *   - Every call returns an int status: 0 success, ALAMO_EL_NOT_CONVERGED (1)
 *     when a solve stopped at max_iter, negative on errors.  Nothing aborts
 *     or throws across this boundary; alamo_el_last_error() describes the
 *     last failure on the calling thread.
 *   - Fields are caller-allocated DEVICE pointers (hipMalloc'd or from any
 *     allocator on the current device) unless a name ends in _host.
 *     One box per level; nodes n = n_cell + 1 per axis; node index i
 *     fastest, then j, then k; vector fields hold 3 components, component
 *     slowest: v[c*N + (k*ny + j)*nx + i].
 *   - Stiffness symmetry tags (template argument SYM of Operator::Elastic,
 *     src/Operator/Elastic.cpp:928-931):
 *       ALAMO_EL_ISOTROPIC  2 values per node  (lambda, mu)
 *       ALAMO_EL_MAJORMINOR 21 values per node (slots of Set::Matrix4<3,MajorMinor>,
 *                           src/Set/Matrix4_MajorMinor.H:327-375)
 *       ALAMO_EL_DIAGONAL   9 values per node  (A[a][b] of Set::Matrix4<3,Diagonal>)
 *     Layout ALAMO_EL_SOA: C[s*N + node]; ALAMO_EL_AOS: C[node*ncomp + s],
 *     the memory order of amrex::BaseFab<Set::Matrix4<3,SYM>>.
 *   - Boundary table: 26 regions x 3 components of type and value, region
 *     order of BC::Operator::Elastic::Elastic::Face
 *     (src/BC/Operator/Elastic/Elastic.H:35-43); types ALAMO_EL_BC_*.
 *   - Every launch goes to the hipStream_t passed in (0 = null stream).
 *     Calls on one handle are not thread safe; distinct handles are.
 */
#ifndef ALAMO_ELASTIC_H
#define ALAMO_ELASTIC_H

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ALAMO_EL_OK 0
#define ALAMO_EL_NOT_CONVERGED 1
#define ALAMO_EL_EINVAL (-1)
#define ALAMO_EL_EHIP (-2)
#define ALAMO_EL_ENAN (-3)
#define ALAMO_EL_ESTATE (-4)

#define ALAMO_EL_ISOTROPIC 0
#define ALAMO_EL_MAJORMINOR 1
#define ALAMO_EL_DIAGONAL 2

#define ALAMO_EL_SOA 0
#define ALAMO_EL_AOS 1

#define ALAMO_EL_BC_DISPLACEMENT 0
#define ALAMO_EL_BC_TRACTION 1
#define ALAMO_EL_BC_PERIODIC 2
#define ALAMO_EL_BC_NEUMANN 3
#define ALAMO_EL_BC_NONE 4

#define ALAMO_EL_PREC_FP64 0      /* plain FP64 multigrid */
#define ALAMO_EL_PREC_FP32_IR 1   /* FP32 cycles inside FP64 iterative refinement */

#define ALAMO_EL_CYCLE_F 0
#define ALAMO_EL_CYCLE_V 1

#define ALAMO_EL_OUTER_MLMG 0     /* x += cycle(b - A x) until converged */
#define ALAMO_EL_OUTER_FGMRES 1   /* flexible GMRES, one cycle as preconditioner */

/* Smoother of the cycles: the reference's damped Jacobi (2 sweeps per Fsmooth
 * call, src/Operator/Operator.cpp:88-154), or red-black Gauss-Seidel (no
 * counterpart in the reference): one Fsmooth call is a red and a black half
 * sweep over the colors (i + j + k) % 2, each x += w (b - A x) / D on its color
 * with w = 0.9 on interior rows and 1/2 on boundary rows. */
#define ALAMO_EL_SMOOTH_JACOBI 0
#define ALAMO_EL_SMOOTH_RBGS 1

#define ALAMO_EL_BOTTOM_BICGSTAB 0 /* the reference bottom solver (Linear.H:258) */
/* The cycles stop at the finest MG level with at most 4096 unknowns and solve
 * there with the FP64 inverse of its operator (computed on the device by
 * alamo_el_prepare); no counterpart in the reference, whose bottom is always
 * the coarsest level. */
#define ALAMO_EL_BOTTOM_DIRECT 1

/* Learned coarse correction (alamo_el_set_learned): the network proposes
 * e = P^L N(R^L r) for the residual r, smoothed by 2 Fsmooth calls; the step
 * x += g e takes the g that minimizes |r - g A e|_2 and is kept only when it
 * lowers |r|_inf (the multigrid residual is the judge). */
#define ALAMO_EL_LEARNED_GUESS 1   /* one learned correction of the initial x */
#define ALAMO_EL_LEARNED_PRECOND 2 /* one between cycles (MLMG) or at each FGMRES restart */

typedef struct alamo_el_s* alamo_el_t;
typedef struct alamo_nn_s* alamo_nn_t; /* alamo_unet.h */

/* MLMG settings; alamo_el_default_opts fills the reference defaults
 * (src/Solver/Nonlocal/Linear.H:255-290 and the amrex::MLMG defaults). */
typedef struct {
    int max_iter;          /* 200 */
    int nu1;               /* 2  Fsmooth calls before restriction (pre_smooth)  */
    int nu2;               /* 2  Fsmooth calls after interpolation (post_smooth) */
    int cycle;             /* ALAMO_EL_CYCLE_F: setMaxFmgIter(1e8), Linear.H:261 */
    int bottom_max_iter;   /* 200 */
    int final_smooth;      /* 8  smooths after a failed bottom solve */
    double omega;          /* 2/3, src/Operator/Operator.H:123 */
    double bottom_tol_rel; /* 1e-4 */
    int precision;         /* ALAMO_EL_PREC_* */
    int outer;             /* ALAMO_EL_OUTER_* */
    int restart;           /* FGMRES restart length (default 30) */
    int verbose;           /* 0 */
    int bottom_solver;     /* ALAMO_EL_BOTTOM_BICGSTAB */
    int learned;           /* 0, or ALAMO_EL_LEARNED_* bits (needs alamo_el_set_learned) */
    int smoother;          /* ALAMO_EL_SMOOTH_JACOBI */
    int max_fmg_iter;      /* -1: every cycle an F-cycle (with cycle F; setMaxFmgIter(1e8),
                              Linear.H:261); k >= 0: F-cycles for the first k iterations,
                              V-cycles after (max_fmg_iter, Linear.H:266,305) */
    int fixed_iter;        /* -1; k > 0: at most k iterations, and a solve that has not
                              converged by then returns 0 with info.res its residual
                              (fixed_iter, "exiting gracefully", Linear.H:267,312) */
} alamo_el_opts;

typedef struct {
    int status;            /* same as the return value of alamo_el_solve */
    int iters;             /* outer iterations (MLMG cycles or FGMRES steps) */
    double res0;           /* |b - A x0|_inf */
    double rhsnorm;        /* |b|_inf */
    double res;            /* final |b - A x|_inf (FP64) */
    double seconds;        /* wall time of the solve on the host clock */
    int learned_steps;     /* learned corrections taken (initial guess included) */
} alamo_el_info;

const char* alamo_el_last_error(void);
void alamo_el_default_opts(alamo_el_opts* opts);

/* Operator::Elastic<SYM>::Elastic + define (src/Operator/Elastic.cpp:10-55):
 * one AMR level, MG hierarchy coarsened by 2 while every axis keeps at least
 * 2 cells, or to max_coarsening_level (-1 = no limit; LPInfo,
 * src/Integrator/Base/Mechanics.H:190-193).  dx = cell size on the finest
 * level.  The handle owns every coefficient, diagonal and work array. */
int alamo_el_create(const int n_cell[3], const double dx[3], int sym, int max_coarsening_level,
                    alamo_el_t* out);
int alamo_el_destroy(alamo_el_t h);
int alamo_el_num_levels(alamo_el_t h, int* nlev);
int alamo_el_level_nodes(alamo_el_t h, int mglev, int nodes[3]);

/* Periodic axes: the geometry's is_periodic (amrex::Geometry; the operator
 * classifies boundary rows on growPeriodicDomain(1),
 * src/Operator/Elastic.cpp:145-149, restriction on the grown coarse domain,
 * src/Operator/Operator.cpp:268, and Constant::Init skips those faces,
 * src/BC/Operator/Elastic/Constant.H:160-185).  On a periodic axis the nodes
 * 0 and n are one point: the nodes of its faces take interior rows with the
 * neighbours across the face wrapped (below node 0 is node n - 1), gradients
 * there are central, and the faces' regions of the boundary table are never
 * used.  Call before alamo_el_set_bc when the table has periodic entries: a
 * periodic entry in a region that lies off the periodic axes is EINVAL.  Not
 * with alamo_el_set_psi, nor on an AMR patch or under one (EINVAL).  With no
 * displacement row left (every axis periodic, or periodic axes and only
 * traction and Neumann rows) the operator is singular. */
int alamo_el_set_periodic(alamo_el_t h, const int periodic[3]);

/* Operator::Elastic::SetBC with a BC::Operator::Elastic::Constant table
 * (src/Operator/Elastic.H:61-65, src/BC/Operator/Elastic/Constant.H). */
int alamo_el_set_bc(alamo_el_t h, const int type_host[26 * 3], const double val_host[26 * 3]);
/* Operator::Elastic::SetUniform (src/Operator/Elastic.H:105). */
int alamo_el_set_uniform(alamo_el_t h, int uniform);
/* Operator<Grid::Node>::SetOmega (src/Operator/Operator.H:57). */
int alamo_el_set_omega(alamo_el_t h, double omega);
/* Operator::Elastic::SetModel(amrlev, FabArray<BaseFab<Matrix4>>): copies the
 * caller's per-node stiffness into the operator's own storage for MG level 0
 * (src/Operator/Elastic.cpp:87-116). */
int alamo_el_set_model(alamo_el_t h, const double* C, int layout, hipStream_t stream);
/* Operator::Elastic::SetPsi(amrlev, psi, psi_small) (src/Operator/Elastic.H:55-57,
 * src/Operator/Elastic.cpp:118-137): psi holds one value per cell of the finest
 * level ((nx-1)(ny-1)(nz-1) device doubles, x fastest; cells outside the domain
 * count as 1).  Every row of Fapply, Diagonal, Fsmooth, the residuals and the
 * solves then carries the weight psi_avg = (1 - psi_small) * (mean of the 8
 * cells around the node) + psi_small, and interior rows gain
 * (1 - psi_small) (C : grad u) . grad psi (Elastic.cpp:190-259, 340-363).  The
 * coarse MG levels take psi averaged down as the reference does
 * (Elastic.cpp:839-907) at the next alamo_el_prepare.  The reference's default
 * psi_small is 1e-8.  A solve with psi needs a coarsest level of at most 1536
 * unknowns or ALAMO_EL_BOTTOM_DIRECT. */
int alamo_el_set_psi(alamo_el_t h, const double* psi, double psi_small, hipStream_t stream);
/* Diagnostic, no counterpart in the reference: the number of rows of MG level
 * mglev's C palette (0 when the level reads its C copy).  A MajorMinor level
 * served by the LDS-tiled kernel whose nodes take at most 65536 distinct
 * stiffness tensors, and at most one per 16 nodes (a polycrystal: one per
 * grain, plus boundary mixtures on its averaged levels), reads a two-byte row
 * index per interior node and the tensor from that table instead of 21 values
 * per node; the products are the same.  ALAMO_EL_PALETTE=0 in the environment at
 * alamo_el_set_model turns it off. */
int alamo_el_palette_size(alamo_el_t h, int mglev, int* npal);
/* prepareForSolve: averageDownCoeffsSameAmrLevel on every MG level and
 * Diagonal(true) (src/Operator/Operator.cpp:253-260,
 * src/Operator/Elastic.cpp:605-626,746-910). */
int alamo_el_prepare(alamo_el_t h, hipStream_t stream);
/* The FP32 copies the FP32 cycles read (C, its tile and boundary copies, the
 * palette table, the compact grad-C matrices and the diagonal of every MG
 * level), after alamo_el_prepare.  No counterpart in the reference, which has
 * no FP32 path; the first ALAMO_EL_PREC_FP32_IR solve after a model change
 * makes them otherwise.  A caller that updates the model every time step
 * calls it with alamo_el_prepare to keep the set-up out of the solve. */
int alamo_el_prepare_fp32(alamo_el_t h, hipStream_t stream);

/* The FP32 kernels of the FP32 cycles, one call each, on MG level mglev with
 * the level's FP32 copies (C, its tile and boundary copies or palette, the
 * compact grad-C matrices, the diagonal; made here if needed).  No
 * counterpart in the reference, which has no FP32 path: for parity tests of
 * those kernels and for callers driving their own FP32 cycles.  op 0: out =
 * A x (Fapply); 1: out = b - A x; 2: one damped Jacobi sweep of x (Fsmooth is
 * two); 3: out = the restriction of x onto level mglev + 1; 4: out += the
 * interpolation of x from level mglev + 1.  Device float arrays of 3 planes. */
int alamo_el_level_op32(alamo_el_t h, int mglev, int op, const float* x, const float* b, float* out,
                        hipStream_t stream);

/* Operator::Elastic::Fapply, f = A u on MG level mglev
 * (src/Operator/Elastic.cpp:139-292). */
int alamo_el_fapply(alamo_el_t h, int mglev, const double* u, double* f, hipStream_t stream);
/* Operator::Elastic::Diagonal (src/Operator/Elastic.cpp:296-379); also cached
 * by alamo_el_prepare. */
int alamo_el_diagonal(alamo_el_t h, int mglev, double* diag, hipStream_t stream);
/* Operator<Grid::Node>::Fsmooth called nsmooth times: 2*nsmooth damped
 * Jacobi sweeps of x against b (src/Operator/Operator.cpp:88-154).  Needs
 * alamo_el_prepare. */
int alamo_el_fsmooth(alamo_el_t h, int mglev, double* x, const double* b, int nsmooth, hipStream_t stream);
/* The red-black Gauss-Seidel smoother (ALAMO_EL_SMOOTH_RBGS) called nsmooth
 * times on MG level mglev: nsmooth red and black half sweeps of x against b.
 * No counterpart in the reference (see alamo_el_fsmooth for its smoother).
 * Needs alamo_el_prepare. */
int alamo_el_gs_smooth(alamo_el_t h, int mglev, double* x, const double* b, int nsmooth, hipStream_t stream);
/* correctionResidual / solutionResidual: r = b - A x
 * (src/Operator/Operator.cpp:635-657). */
int alamo_el_residual(alamo_el_t h, int mglev, const double* x, const double* b, double* r, hipStream_t stream);
/* Operator<Grid::Node>::restriction: crse (level cmglev) = R fine (level
 * cmglev-1) (src/Operator/Operator.cpp:262-363). */
int alamo_el_restrict(alamo_el_t h, int cmglev, double* crse, const double* fine, hipStream_t stream);
/* Operator<Grid::Node>::interpolation: fine (level fmglev) += P crse
 * (src/Operator/Operator.cpp:365-434). */
int alamo_el_interp_add(alamo_el_t h, int fmglev, double* fine, const double* crse, hipStream_t stream);

/* BC::Operator::Elastic::Constant::Init on the finest level: boundary nodes
 * of rhs get the boundary values, or 0 for displacement components when
 * homogeneous (src/BC/Operator/Elastic/Constant.H:197-286). */
int alamo_el_bc_init(alamo_el_t h, double* rhs, int homogeneous, hipStream_t stream);
/* Base::Mechanics post-processing on the finest level, small-strain models
 * (src/Integrator/Base/Mechanics.H:210-245): sigma = DW(grad u) =
 * C:(grad u - F0) and, when strain is not NULL, strain = (grad u + grad u^T)/2,
 * at every node, grad u central inside and one-sided on the domain faces.
 * sigma and strain: 9 planes (3p+q) of N; C in the layout passed; F0 NULL or
 * 9 planes. */
int alamo_el_stress(alamo_el_t h, const double* C, int layout, const double* F0, const double* u, double* sigma,
                    double* strain, hipStream_t stream);
/* Base::Mechanics::Integrate, 3D (src/Integrator/Base/Mechanics.H:391-448):
 * the traction on the domain face normal to axis (0..2) at side (0 low,
 * 1 high), sum over the face's cells of 0.25 (sigma at the 4 corner nodes) . da
 * with da = e_axis * (product of the other two cell sizes), and the
 * displacement at the face node whose two in-face indices are n - 2 (the
 * reference's trac_hi / disp_hi of thermo.dat for axis 0, side 1).  Synchronizes
 * the stream; trac_host and disp_host (may be NULL) receive 3 values. */
int alamo_el_face_traction(alamo_el_t h, const double* sigma, const double* u, int axis, int side,
                           double* trac_host, double* disp_host, hipStream_t stream);
/* Integrator::Mechanics::UpdateModel on the finest level
 * (src/Integrator/Mechanics.H:249-253): out[c*N + node] =
 * sum_n eta[n*N + node] * table_host[n*ncomp + c], from zero in model order,
 * each product rounded before it is added, like the reference's
 * `model += eta(i,j,k,n) * models[n]`.  eta and out are device fields
 * (nmodels and ncomp planes); table_host holds each model's components on the
 * host: its C in the handle's slot order (2 Isotropic, 21 MajorMinor, 9
 * Diagonal; ALAMO_EL_SOA), or its F0 (9 values, 3a+b).  nmodels * ncomp <= 256. */
int alamo_el_mix_models(alamo_el_t h, const double* eta, int nmodels, const double* table_host, int ncomp,
                        double* out, hipStream_t stream);
/* Newton<T>::prepareForSolve right-hand side, psi unset
 * (src/Solver/Nonlocal/Newton.H:151-231): rhs = b - BC(u, grad u, DW) on the
 * boundary and b - div DW in the interior, with DW = C:(grad u - F0).  F0 is
 * NULL or 9 values per node (F0[(3a+b)*N + node]); C uses the layout passed. */
int alamo_el_newton_rhs(alamo_el_t h, const double* C, int layout, const double* F0, const double* u,
                        const double* b, double* rhs, hipStream_t stream);

/* alamo_el_newton_rhs without allocation.  No counterpart in the reference:
 * it is for callers that cannot call hipMalloc (code under HIP graph capture,
 * time-stepping or serving loops, harnesses that forbid allocation).  Same
 * arguments and result as alamo_el_newton_rhs (bit for bit); instead of the
 * handle's on-demand 18 N doubles it marches the box in slabs of z planes
 * through a fixed 32 MiB device array of the library (one per GPU, held until
 * the process ends), so a plane may hold at most 155,344 nodes (e.g. 394 x 394);
 * larger planes return ALAMO_EL_EINVAL.  Calls that overlap on one GPU (other
 * streams or threads) share that array and race. */
int alamo_el_newton_rhs_noalloc(alamo_el_t h, const double* C, int layout, const double* F0, const double* u,
                                const double* b, double* rhs, hipStream_t stream);

/* Solver::Nonlocal::Linear::solve(sol, rhs, tol_rel, tol_abs)
 * (src/Solver/Nonlocal/Linear.H:131-147): solves A x = b on the finest level
 * to |b - A x|_inf <= max(tol_abs, tol_rel * max(|b|_inf, |b - A x0|_inf)).
 * x holds the initial guess on entry.  opts may be NULL (defaults).  Needs
 * alamo_el_set_model; calls alamo_el_prepare when the model changed. */
int alamo_el_solve(alamo_el_t h, double* x, const double* b, double tol_rel, double tol_abs,
                   const alamo_el_opts* opts, alamo_el_info* info, hipStream_t stream);

/* |v|_inf over 3 components of a level-mglev vector, on the host. */
int alamo_el_norm_inf(alamo_el_t h, int mglev, const double* v, double* out_host, hipStream_t stream);

/* --- Learned coarse-level solver.  No counterpart in the reference: these
 * connect the network of alamo_unet.h to the multigrid hierarchy. --- */
/* Copy of the stiffness of MG level mglev (SoA, as set by alamo_el_set_model
 * and averaged down by alamo_el_prepare). */
int alamo_el_level_model(alamo_el_t h, int mglev, double* C, hipStream_t stream);
/* Level-mglev right-hand side of a finest-level b: interior rows restricted
 * by the MG full weighting level by level, boundary rows injected. */
int alamo_el_coarse_rhs(alamo_el_t h, int mglev, const double* b, double* bc, hipStream_t stream);
/* Network input on level mglev from its right-hand side bc: float32
 * [nc + 30][N]: C / alpha (alpha = mean of slot 0), interior bc / (alpha s),
 * boundary bc / s, displacement-row indicators, node coordinates scaled to
 * [0, 1], and 18 constant channels holding the mean of bc / s over each face
 * (xlo, ylo, zlo, xhi, yhi, zhi) and component; *scale_host = s (the largest
 * scaled magnitude, 0 for bc = 0).  The network predicts u / s. */
int alamo_el_learned_input(alamo_el_t h, int mglev, const double* bc, float* features, double* scale_host,
                           hipStream_t stream);
/* Attach a network (NULL detaches) working on MG level mglev: its grid is
 * that level's nodes, nc + 30 inputs, 3 outputs.  The handle does not own it. */
int alamo_el_set_learned(alamo_el_t h, alamo_nn_t nn, int mglev);
/* x = P^L (s * net(input(coarse_rhs(b)))) on the finest level. */
int alamo_el_learned_guess(alamo_el_t h, const double* b, double* x, hipStream_t stream);

/* --- Composite AMR solve (refinement ratio 2, one box per AMR level):
 * Operator::reflux, fixUpResidualMask and averageDownSolutionRHS
 * (src/Operator/Operator.cpp:230-251, 436-448, 526-633) with MLMG's sweep over
 * AMR levels. --- */
/* Integrator::Mechanics::TagCellsForRefinement on the finest level of h
 * (src/Integrator/Mechanics.H:298-313): cell (i, j, k) is tagged when
 * |grad eta_n| |dx| > threshold (eta_ref_threshold, 0.01 by default in the
 * reference, src/Integrator/Mechanics.H:137) at node (i, j, k) for some component n of eta
 * (ncomp device planes of the level's nodes; the gradient central inside and
 * one-sided on the box faces).  *ntag receives the number of tagged cells and,
 * when it is not 0, [clo, chi] the tagged cells' bounding box grown by buffer
 * cells and clipped to the level, in the level's cell indices (the one box per
 * level that alamo_amr_patch_create takes; the clustering of tags into several
 * boxes by AmrCore is not restated).  Synchronizes the stream. */
int alamo_el_tag_bbox(alamo_el_t h, const double* eta, int ncomp, double threshold, int buffer, int clo[3],
                      int chi[3], long long* ntag, hipStream_t stream);
/* The same tags, one byte per cell of h's finest level (device array of
 * (nx-1)(ny-1)(nz-1) bytes, x fastest; 1 = tagged). */
int alamo_el_tag_cells(alamo_el_t h, const double* eta, int ncomp, double threshold, unsigned char* tags,
                       hipStream_t stream);
/* Boxes of the next AMR level from cell tags (AmrCore::MakeNewGrids with
 * amrex::ClusterList; parameters amr.n_error_buf, amr.blocking_factor,
 * amr.grid_eff, src/Integrator/Integrator.cpp:23-140).  Host only, no GPU work.
 * tags_host: ncell[0] x ncell[1] x ncell[2] bytes (x fastest) whose cell
 * (0, 0, 0) is global cell origin.  Tags are grown by buffer cells, coarsened
 * by blocking_factor / 2 on the global grid (boxes aligned to it), clustered by
 * Berger-Rigoutsos (a box is kept when its tagged fraction reaches grid_eff,
 * else cut at a hole of the tag signature, else at its strongest inflection,
 * else in the middle of its longest axis), refined back and clipped to the
 * array; boxes closer than one cell are merged, so that the patches of a level
 * never share nodes (the composite solve couples a patch with its parent
 * only).  boxes receives 6 ints per box: clo[3], chi[3] in global cells, the
 * form alamo_amr_patch_create takes (chi = last cell + 1). */
int alamo_amr_cluster(const unsigned char* tags_host, const int ncell[3], const int origin[3], int buffer,
                      int blocking_factor, double grid_eff, int max_boxes, int* boxes, int* nboxes);
/* Patch of the next AMR level covering the parent's cells [clo, chi] (global
 * parent-level cell indices, clo < chi <= domain cells).  The new handle holds
 * the fine nodes 2 clo .. 2 chi plus one ghost node layer on each face inside
 * the domain (values interpolated from the parent, the "ghostnodes" treatment);
 * its boundary table is the parent's on domain faces and displacement on ghost
 * faces.  The patch and its ghost layer must lie on the parent's unknowns.  Set
 * its model with alamo_el_set_model over the handle's nodes
 * (alamo_amr_patch_info gives the global node index of node 0). */
int alamo_amr_patch_create(alamo_el_t parent, const int clo[3], const int chi[3], alamo_el_t* out);
/* Elastic::averageDownCoeffsDifferentAmrLevels (src/Operator/Elastic.cpp:628-745;
 * Linear's average_down_coeffs, Linear.H:288): the parent's stiffness at its
 * nodes clo..chi (covered and coarse-fine nodes) becomes the full-weighting
 * restriction of the patch's, with the corner, edge and face weights on domain
 * faces.  patch_C: the patch's C over all its nodes (ghost layer included),
 * parent_C: the parent's, both SoA in the handles' symmetry (device); the
 * caller then sets the parent's model from parent_C (alamo_el_set_model), as
 * the reference's prepareForSolve does before averaging the MG levels.  Call
 * it for the deepest patches first, as the reference loops amrlev down. */
int alamo_amr_average_down_coeffs(alamo_el_t patch, const double* patch_C, double* parent_C, hipStream_t stream);
int alamo_amr_patch_info(alamo_el_t h, int origin[3], int ghost_lo[3], int ghost_hi[3]);
/* Regrid: fill the field dst_u of the new patch dst (every handle node, ghost
 * layer included) as Integrator::FillPatch does for a nodal field
 * (src/Integrator/Integrator.cpp:394-446, FillPatchTwoLevels with
 * node_bilinear_interp): nodes that the old patch src of the same AMR level
 * holds as unknowns (its global nodes 2 clo .. 2 chi) take src_u's value, the
 * others the trilinear interpolation of parent_u, the field of dst's parent
 * handle.  With src NULL every node is interpolated, as
 * Integrator::FillCoarsePatch does for a new level (:452-478).  The old patch
 * may have another parent (the regridded one). */
int alamo_amr_fillpatch(alamo_el_t dst, const double* parent_u, alamo_el_t src, const double* src_u, double* dst_u,
                        hipStream_t stream);
/* Operator::reflux (src/Operator/Operator.cpp:526-633) of one patch into its
 * parent: parent_res (a field of the parent handle) at the parent nodes
 * clo..chi becomes the nodal full weighting of fine_res (a field of the patch
 * handle; its ghost-layer values are taken as 0), with the face, edge and
 * corner rules of the MG restriction on domain faces.  fine_res is not
 * modified. */
int alamo_amr_reflux(alamo_el_t patch, const double* fine_res, double* parent_res, hipStream_t stream);
/* Gather or scatter one box of a multi-box field (an amrex::MultiFab's FABs on
 * one AMR level) to or from the single-box layout of a level: copies the nodes
 * [lo, hi] (global node indices, inclusive) of ncomp planes from src (src_n
 * nodes per axis, its node 0 at global src_origin) into dst (dst_n, dst_origin).
 * The box must lie inside both arrays. */
int alamo_el_box_copy(double* dst, const int dst_n[3], const int dst_origin[3], const double* src, const int src_n[3],
                      const int src_origin[3], const int lo[3], const int hi[3], int ncomp, hipStream_t stream);
/* Solve the composite system over levels[0] (the base handle) and its nested
 * patches levels[1..nlev-1], each patch's parent listed before it (a level may
 * hold several patches that share no nodes, as alamo_amr_cluster makes them):
 * x[l], b[l] are fields of each handle.  One
 * iteration pre-smooths every patch (nu1 Fsmooth calls), refluxes the residual
 * into its parent, runs one cycle (opts->cycle, opts->precision) on the base
 * level, interpolates the corrections up and post-smooths (nu2); then the
 * solution is averaged down and the ghost layers refreshed.  Stops on the
 * composite inf-norm (parent nodes strictly covered by a patch excluded,
 * coarse-fine nodes included). */
int alamo_amr_solve(int nlev, const alamo_el_t* levels, double* const* x, const double* const* b, double tol_rel,
                    double tol_abs, const alamo_el_opts* opts, alamo_el_info* info, hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif
