/*
 * CPU oracle for the MLFMAxwell hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A single-threaded C restatement of the reference's MLFMA operator setup and
 * matvec (modMLFMA.f90), ILU0/CGS (modIterativo.f90), plane-wave excitation
 * (main.f90 generar_onda) and RCS (modRCS.f90).  Only tests/, the smoke()
 * check and the CPU-baseline leg of bench.py may load it; the product path
 * (lywell_amd/) never links or calls it.
 *
 * Conventions follow the reference's Fortran arrays so that the same buffers
 * can be handed to the oracle and to the product C-ABI:
 *   - geometry arrays are column-major (3,n) etc., i.e. xyz contiguous;
 *   - integer connectivity holds 1-based indices;
 *   - complex vectors are interleaved (re, im) doubles.
 */
#ifndef LYWELL_ORACLE_H
#define LYWELL_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_mlfma or_mlfma;

/* modRWG.f90:8-174 + 270-328: RWG edges from a triangle list.
 * t_p (3,nt) 1-based.  Outputs (caller-allocated, capacity 3*nt edges):
 * e_p(2,ne), e_t(2,ne), e_po(2,ne), e_long(ne), e_centro(3,ne); returns ne. */
int or_rwg_build(int np, const double *p_coord, int nt, const int *t_p,
                 int *e_p, int *e_t, int *e_po, double *e_long, double *e_centro);
/* hallar_geometria_rwg: t_area(nt), t_baric(3,nt), t_baric_sub(3,9,nt). */
void or_tri_geometry(int np, const double *p_coord, int nt, const int *t_p,
                     double *t_area, double *t_baric, double *t_baric_sub);

/* inicializar_MLFMA (modMLFMA.f90:113-275).  fp64_near=0 keeps the
 * reference's complex(4) near-field storage (dp_near=4). Returns NULL and sets
 * *status on failure (1 = octree too small, the reference's STOP). */
or_mlfma *or_mlfma_init(double lambda, double ladomax, int fp64_near,
                        int ne, const double *e_centro, const int *e_t,
                        const int *e_po, const double *e_long, int nt,
                        const double *t_area, const double *t_baric,
                        const double *t_baric_sub, int np,
                        const double *p_coord, const int *t_p, const int *e_p,
                        const double *t_normal, int *status);
void or_mlfma_free(or_mlfma *h);

/* Structure queries (levels 1..L, 1-based as in the reference). */
int or_nlevels(const or_mlfma *h);
int or_level_info(const or_mlfma *h, int level, int *nboxes, int *nmult,
                  int *nth, int *nph, int *ntransf, long long *npairs);
long long or_near_nnz(const or_mlfma *h);
/* Near CSR view: ia(ne+1), ja(nnz) 1-based; z as interleaved doubles. */
void or_near_csr(const or_mlfma *h, int *ia, int *ja, double *z);
/* Copy a per-level table: which=0 kappas(3,K), 1 ang(2,K), 2 pesos(K),
 * 3 mat_interp weights (9K), 4 mat_interp indices (9K, as doubles),
 * 5 functransf (K,nT) interleaved complex, 6 funcRAD (K,2,ne) complex,
 * 7 transfer vectors rab (3,nT) (observer box centre minus source box centre). */
long long or_level_table(const or_mlfma *h, int level, int which, double *out);
/* Leaf box of each basis (1-based), box centres and parents per level. */
void or_leaf_of_basis(const or_mlfma *h, int *leaf);
void or_box_centres(const or_mlfma *h, int level, double *c, int *parent);
/* Far list of every box at a level, in list order (cubLej_indFuncTransf(1,:),
 * modMLFMA.f90:1840-1841): ptr (nboxes + 1) and 0-based far boxes. Returns
 * the total count; with ptr == NULL only the count. */
long long or_far_lists(const or_mlfma *h, int level, int *ptr, int *boxes);

/* MatxVec_MLFMA (modMLFMA.f90:1901-1946). x, y: interleaved complex[ne]. */
void or_matvec(or_mlfma *h, const double *x, double *y);
/* Far-only / near-only parts of the same product (for per-stage checks). */
void or_matvec_far(or_mlfma *h, const double *x, double *y);
void or_matvec_near(or_mlfma *h, const double *x, double *y);
/* Zmn_MoM (modMoM.f90:89-135), full FP64, 1-based m, n. */
void or_zmn(or_mlfma *h, int m, int n, double *z);

/* ilu0C (modIterativo.f90:1805-1976) on a CSR (1-based ia/ja).  Storage of the
 * factor is complex(4) when fp64=0, complex(8) otherwise; alu is returned as
 * interleaved doubles either way.  alu/jlu sized nnz+1, ju sized n.
 * Returns ierr (0 ok, k = zero pivot at row k). */
int or_ilu0(int n, const int *ia, const int *ja, const double *a, int fp64,
            double *alu, int *jlu, int *ju);
/* lusol0C (modIterativo.f90:504-560). */
void or_lusol0(int n, const double *y, double *x, const double *alu,
               const int *jlu, const int *ju);

/* CGSold (modIterativo.f90:331-493) with MatxVec_MLFMA, optional ILU0.
 * maxit caps the loop (the reference has none); returns the iteration count,
 * *converged = 1 when ||r|| < errmax ||b||, *relres = ||r||/||b|| (recurrence). */
int or_cgs(or_mlfma *h, const double *b, const double *guess, double errmax,
           const double *alu, const int *jlu, const int *ju, int maxit,
           double *x, int *converged, double *relres);

/* generar_onda (main.f90:1488-1525): b[ne] for direction d, complex ctte and
 * complex polarisation pol(3) (interleaved). */
void or_plane_wave(or_mlfma *h, const double *d, const double *ctte,
                   const double *pol, double *b);
/* sigmaRCS (modRCS.f90:91-153) for current vector I; returns m^2. */
double or_sigma_rcs(or_mlfma *h, const double *I, const double *pol,
                    const double *ctte, double theta, double phi);

/* Numerical kernels exposed for closed-form checks. */
double or_legendre(int n, double x);            /* modGlobalMetod.f90:988 */
void or_hankel(int n, double x, double *h);       /* :493, h = j_n - i y_n */
double or_norm2(const double *v);                 /* flang NORM2 semantics */
void or_set_lambda(double lambda);                /* modGlobalParam setlambda */
double or_get_w(void);

#ifdef __cplusplus
}
#endif
#endif
