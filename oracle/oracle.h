/*
 * TEST INFRASTRUCTURE ONLY — the oracle.
 *
 * CPU restatement (plain C + OpenMP, FP64) of TopOpt.jl v0.14.0's SIMP
 * compliance path: structured PointLoadCantilever mesh and Ferrite DOF
 * numbering, element stiffness by Gauss quadrature, the matrix-free operator,
 * IterativeSolvers-style Jacobi PCG, element energy / compliance sensitivity,
 * the BFS-built DensityFilterFun and its transpose, volume, and the
 * build-defined OC update.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker / CPU baseline.  The
 * product (libtopopt_hip) never links or calls it.
 *
 * Every function cites the reference file:line it restates.  Conventions:
 * arrays indexed like the reference (Ferrite DOF order, 1-based integer
 * payloads in node_dofs / cell_dofs, column-major matrices).
 */
#ifndef TOPOPT_ORACLE_H
#define TOPOPT_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Mesh sizes of a dim-D structured grid with nels[] cells. */
int64_t orc_nnodes(int dim, const int64_t* nels);
int64_t orc_ncells(int dim, const int64_t* nels);

/* Ferrite cell-first-touch DOF numbering (Metadata node_dofs / cell_dofs).
 * node_dofs: dim x nnodes, column-major, 1-based.  cell_dofs: ndpc x nel,
 * column-major, 1-based.  cell_nodes: nnpc x nel, 1-based node ids
 * (x-fastest).  Returns ndofs. */
int64_t orc_numbering(int dim, const int64_t* nels, int64_t* node_dofs,
                      int64_t* cell_dofs, int64_t* cell_nodes);

/* Element stiffness (Ferrite local order dim*a+d, row-major = column-major
 * since symmetric), isoparametric Q4/hex8 on an h-sized box, quad_order Gauss
 * points per axis (2..4), plane strain in 2D. */
int orc_element_stiffness(int dim, double E, double nu, const double* h,
                          int quad_order, double* Ke);

/* Penalised density rho_e and d rho_e / d x_e (power penalty). */
void orc_rho_drho(double x, double p, double xmin, int pbi, double* rho,
                  double* drho);
/* Same for penalty family kind (0 power, 1 rational, 2 sinh) after
 * projection proj (0 none, 1 Heaviside, 2 sigmoid) of steepness beta. */
void orc_rho_drho_fun(double x, int kind, double p, int proj, double beta, double xmin,
                      int pbi, double* rho, double* drho);
/* Projection proj (0 none, 1 Heaviside, 2 sigmoid) value and derivative. */
void orc_project(double x, int proj, double beta, double* v, double* dv);

/* Operator context (restates MatrixFreeOperator fields). */
typedef struct orc_op {
    int dim, ndpc;
    int64_t nel, ndof;
    const int64_t* cell_dofs; /* 1-based, ndpc x nel */
    const double* Ke;         /* ndpc x ndpc */
    const uint8_t* fixed;     /* ndof */
    double xmin, p;
    int pbi;                  /* penalty before interpolation */
    double meandiag;          /* sum_e tr(Ke) */
    /* built by orc_op_init: dof_cells CSR (ascending cell) */
    int64_t* dc_off;  /* ndof+1 */
    int64_t* dc_val;  /* pairs (cell, local) flattened: 2*nnz */
    double* xes;      /* nel x ndpc scratch */
    double* px;       /* nel */
    int kind, proj;   /* penalty family and projection (orc_rho_drho_fun) */
    double beta;      /* projection steepness */
} orc_op;

int orc_op_init(orc_op* op);
void orc_op_free(orc_op* op);
void orc_op_set_vars(orc_op* op, const double* vars);
void orc_matvec(const orc_op* op, const double* x, double* y);
void orc_jacobi_diag(const orc_op* op, double* d);

/* Preconditioned CG with the IterativeSolvers.cg! update order. Returns the
 * iteration count; *resid = final ||r||_2. x is the initial guess when
 * initially_zero == 0. */
int64_t orc_pcg(const orc_op* op, const double* b, double* x, double abstol,
                double reltol, int64_t maxiter, int initially_zero,
                int use_jacobi, double* resid);

/* PCG with the EnergyCriteria stopping test (state in *energy, in/out);
 * *status: 1 converged, 0 maxiter, -1 NaN, -2 negative x'Ax. */
int64_t orc_pcg_energy(const orc_op* op, const double* b, double* x, double tol,
                       int64_t maxiter, int initially_zero, int use_jacobi, double* energy,
                       int* status, double* resid);

/* compute_element_energy: returns obj; ce, grad length nel. */
double orc_element_energy(const orc_op* op, const double* u,
                          const double* vars, double* ce, double* grad);

/* DensityFilterFun Jacobian built exactly like get_neighbour_info + getJacobian
 * (BFS from each cell centre, weights max(rmin-d,0), node<-cell volume
 * average).  CSR by rows (cells), columns ascending.  Returns nnz, or -1 if
 * the filter is empty (the reference throws ArgumentError). Call with
 * rowptr == NULL to get nnz only. */
int64_t orc_filter_build(int dim, const int64_t* nels, const double* h,
                         double rmin, int64_t* rowptr, int64_t* col,
                         double* val);
void orc_csr_mv(int64_t nrows, const int64_t* rowptr, const int64_t* col,
                const double* val, const double* x, double* y);
void orc_csr_mtv(int64_t nrows, int64_t ncols, const int64_t* rowptr,
                 const int64_t* col, const double* val, const double* x,
                 double* y);

/* OC update (build-defined, see oc.c header). Returns lambda. */
/* Cell volume by Gauss quadrature of det J (get_cell_volumes,
 * src/TopOptProblems/elementinfo.jl:180-192) and VolumeFun's value
 * sum_e x_e V_e (/ sum_e V_e when fraction; src/Functions/volume.jl:59-80). */
double orc_cell_volume(int dim, const double* h, int quad_order);
int orc_set_threads(int n);
double orc_volume(int64_t nel, const double* x, const double* cellvol, int fraction);

double orc_oc_update(int64_t nel, const double* x, const double* dc,
                     const double* colsum, double volfrac, double move,
                     double tol, double* xnew);

/* Banded Cholesky direct solve of the BC-reduced assembled K (2D oracle).
 * Returns 0 on success, 1 if not positive definite (reference writes NaN). */
int orc_direct_solve(const orc_op* op, const double* f, double* u);

#ifdef __cplusplus
}
#endif
#endif
