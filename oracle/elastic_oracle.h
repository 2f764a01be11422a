/*
 * elastic_oracle.h -- FP64 CPU restatement of Alamo's nodal elastic
 * multigrid path.  TEST INFRASTRUCTURE ONLY: this library is the checker
 * for the HIP implementation (tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py).  The product library never links it.
 *
 * Every function cites the reference file:line it restates.  Fields use the
 * same conventions as the product C-ABI (include/alamo_elastic.h):
 *   - one box per level, nodes n[0] x n[1] x n[2] (= n_cell + 1 per axis),
 *     node index i fastest, then j, then k;
 *   - vector fields: 3 components, component slowest: v[c*N + idx];
 *   - stiffness fields C: structure-of-arrays, C[s*N + idx], with
 *       sym 0 Isotropic  : s = 0 lambda, 1 mu
 *       sym 1 MajorMinor : s = 0..20, the 21 slots of Set::Matrix4<3,MajorMinor>
 *       sym 2 Diagonal   : s = 3a+b, A[a][b]
 *   - boundary table: 26 regions x 3 components of (type, value) in the
 *     order of BC::Operator::Elastic::Elastic::Face.
 */
#ifndef ALAMO_ELASTIC_ORACLE_H
#define ALAMO_ELASTIC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

int ora_ncomp_of_sym(int sym);

/* The periodic axes of every later call (amrex::Geometry is_periodic; the
 * operator then classifies on growPeriodicDomain, src/Operator/Elastic.cpp:
 * 145-149, src/Operator/Operator.cpp:92,268): on a periodic axis the nodes 0
 * and n-1 are one point, indices past a face wrap to the periodic image, and
 * the faces carry interior rows.  Process-wide; 0 0 0 restores the default. */
int ora_set_periodic(const int per[3]);

/* Operator::Elastic<SYM>::Fapply  (src/Operator/Elastic.cpp:139-292) */
int ora_fapply(const int n[3], const double dx[3], int sym, const double* C,
               const int* bc_type, int uniform, const double* u, double* f);
/* Operator::Elastic<SYM>::Diagonal (src/Operator/Elastic.cpp:296-379) */
int ora_diagonal(const int n[3], const double dx[3], int sym, const double* C,
                 const int* bc_type, double* diag);
/* Operator<Grid::Node>::Fsmooth, nsmooth calls of 2 Jacobi sweeps each
 * (src/Operator/Operator.cpp:88-154) */
int ora_fsmooth(const int n[3], const double dx[3], int sym, const double* C,
                const int* bc_type, int uniform, const double* diag,
                double omega, int nsmooth, double* x, const double* b);
/* r = b - A x (src/Operator/Operator.cpp:635-657) */
int ora_residual(const int n[3], const double dx[3], int sym, const double* C,
                 const int* bc_type, int uniform, const double* x,
                 const double* b, double* r);
/* Operator<Grid::Node>::restriction (src/Operator/Operator.cpp:262-363);
 * nf = fine node counts, crse gets (nf-1)/2+1 nodes per axis */
int ora_restrict(const int nf[3], int ncomp, const double* fine, double* crse);
/* Operator<Grid::Node>::interpolation: fine += P crse
 * (src/Operator/Operator.cpp:365-434) */
int ora_interp_add(const int nf[3], int ncomp, const double* crse, double* fine);
/* Elastic::averageDownCoeffsSameAmrLevel, one level
 * (src/Operator/Elastic.cpp:746-910) */
int ora_average_down_C(const int nf[3], int sym, const double* Cf, double* Cc);
/* BC::Operator::Elastic::Constant::Init on a vector MultiFab
 * (src/BC/Operator/Elastic/Constant.H:197-286) */
int ora_bc_init(const int n[3], const int* bc_type, const double* bc_val,
                int homogeneous, double* rhs);
/* Newton<T>::prepareForSolve, vector version, psi unset
 * (src/Solver/Nonlocal/Newton.H:144-231): rhs = b - N(u).  F0 may be NULL
 * (Linear models); otherwise 9 comps per node, F0[(3a+b)*N+idx]. */
int ora_newton_rhs(const int n[3], const double dx[3], int sym, const double* C,
                   const double* F0, const int* bc_type, const double* u,
                   const double* b, double* rhs);
/* Stress sigma = DW(grad u) with one-sided stencils at the domain faces
 * (src/Integrator/Base/Mechanics.H:210-248); 9 comps per node */
int ora_stress(const int n[3], const double dx[3], int sym, const double* C,
               const double* F0, const double* u, double* sigma);

/* MLMG solve of A x = b on one AMR level (src/Solver/Nonlocal/Linear.H:131-168
 * with PrepareMLMG :255-290 driving amrex::MLMG).  x is the initial guess on
 * entry and the solution on exit.  Returns 0 on convergence, 1 when max_iter
 * was reached, negative on argument errors.
 *   opts[0] max_iter (200)        opts[1] nu1 Fsmooth calls (2)
 *   opts[2] nu2 (2)               opts[3] cycle: 0 F-cycle every iteration
 *                                          (max_fmg_iter=1e8), 1 V-cycle
 *   opts[4] bottom_max_iter (200) opts[5] final_smooth nuf (8)
 *   opts[6] max_coarsening_level (-1 = until 2 cells)
 *   opts[7] F-cycle residual transfer: 1 nodal injection (amrex::average_down,
 *           the MLLinOp::avgDownResMG default), 0 full-weighting restriction
 *   opts[8] max_fmg_iter (-1: F-cycles throughout; k: F for the first k
 *           iterations, V after)      opts[9] fixed_iter (-1; k > 0: at most
 *           k iterations, not converged by then returns 0)
 *   ropts[0] tol_rel  ropts[1] tol_abs  ropts[2] omega (2/3)
 *   ropts[3] bottom_tol_rel (1e-4)
 * out: out_iters, out_res[0] final residual inf-norm, out_res[1] initial,
 * out_res[2] rhs inf-norm, out_hist (may be NULL, max_iter+1 entries). */
int ora_mlmg_solve(const int n[3], const double dx[3], int sym, const double* C,
                   const int* bc_type, int uniform, const int* opts,
                   const double* ropts, double* x, const double* b,
                   int* out_iters, double* out_res, double* out_hist);

/* Number of OpenMP threads the oracle runs on. */
int ora_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
