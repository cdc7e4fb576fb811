/*
 * reaxff_oracle.h — CPU FP64 oracle for the per-step neighbour-list,
 * bond-order and QEq path of the reference (SW_REAXFF).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only the test
 * suite, the smoke check and the CPU-baseline leg of the benchmark load it.
 * The MI355X product never links, loads or calls it.
 *
 * Each stage restates the reference's *live* algorithm (the Sunway offload
 * kernels and the host fallbacks they replaced), in plain C++ with libm
 * exp/pow/log/sqrt, built with -ffp-contract=off.  Per-stage citations are in
 * reaxff_oracle.cpp.
 *
 * Atom indexing follows LAMMPS: atoms [0,n) are local, [n,N) are ghosts
 * (periodic images) whose owner[] is the local atom they copy.  Types passed
 * in are LAMMPS types 1..ntypes; the element map is set at create time like
 * `pair_coeff * * ffield.reax <elements...>`.
 */
#ifndef REAXFF_ORACLE_H
#define REAXFF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_sys orc_sys;

/* Creates a system: parses the force field, maps the `nelem` LAMMPS types to
 * the named elements, and sets the pair_style/fix parameters.  Returns NULL
 * and fills errbuf on failure. */
orc_sys *orc_create(const char *ffield_path, int nelem, const char *const *elements,
                    double bond_cut, double hbond_cut,
                    double swa, double swb, double tol,
                    char *errbuf, int errlen);
void orc_destroy(orc_sys *s);
void orc_set_threads(int nthreads);

/* Force-field and control queries (known-answer tests). */
double orc_gp(orc_sys *s, int k);
int    orc_sbp(orc_sys *s, int elem, double *out16);   /* see field order in .cpp */
int    orc_tbp(orc_sys *s, int ei, int ej, double *out14);
double orc_control(orc_sys *s, const char *key);
void   orc_taper(orc_sys *s, double *tap8);

/* Atoms: n local, N total; x is [N][3]. */
int orc_set_atoms(orc_sys *s, int n, int N, const int *tag, const int *type,
                  const int *owner, const double *x, const double *q);
int orc_set_q(orc_sys *s, const double *q);

/* a1: full Verlet list over all N atoms (local and ghost), d^2 <= cut^2. */
int64_t orc_build_verlet(orc_sys *s, double cutneigh, int brute_force);
int     orc_set_verlet(orc_sys *s, const int64_t *offsets, const int *list);

/* a2/a3: atom pack and far list (d^2 <= nonb_cut^2). Returns entries. */
int     orc_write_reax_atoms(orc_sys *s);
int64_t orc_write_reax_lists(orc_sys *s);

/* a4: capacities and Reset. */
int orc_estimate_storages(orc_sys *s);
int orc_reset(orc_sys *s);

/* a5-a7: bonds + hbonds, list validation. */
int orc_init_forces(orc_sys *s);
int orc_validate_lists(orc_sys *s);

/* a8-a10: BO sections 1-3. */
int orc_bo(orc_sys *s);

/* a11-a16: QEq. */
int orc_qeq_init_matvec(orc_sys *s);
int orc_qeq_cg(orc_sys *s, int which /* 0: s, 1: t */);
/* CG_v2's iteration cap imax (200 in the reference): the solve returns
 * x after imax - 1 updates when it has not converged before. */
void orc_set_cg_imax(orc_sys *s, int imax);
int orc_qeq_calculate_q(orc_sys *s);
int orc_qeq_pre_force(orc_sys *s, int *iters_s, int *iters_t);
int orc_sparse_matvec(orc_sys *s, const double *x_N, double *b_n);

/* Whole per-step path in the reference's order: QEq pre_force, then the pair
 * stages (pack, Reset, far list, Init_Forces, Validate_Lists, BO).  The first
 * call runs the setup-time capacity estimate. */
int orc_step(orc_sys *s, int *iters_s, int *iters_t);

/* Named array access: returns byte size (or -1), copies into out if non-NULL.
 * set_array writes (for capacities, histories). */
int64_t orc_get_array(orc_sys *s, const char *name, void *out);
int     orc_set_array(orc_sys *s, const char *name, const void *in, int64_t bytes);
int64_t orc_get_scalar_i(orc_sys *s, const char *name);

#ifdef __cplusplus
}
#endif

#endif
