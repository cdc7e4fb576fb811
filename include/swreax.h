/*
 * swreax.h — C-ABI of the MI355X (gfx950) neighbour-list + bond-order + QEq
 * path.  Library: sw_reaxff_amd/libswreax.so.
 *
 * The entry points below replace, one for one, the reference's Sunway
 * offload functions on this path (the `_C` functions that a LAMMPS style
 * calls with a POD "pack").  Each comment names the reference function and
 * its caller.  Differences in form, all deliberate:
 *
 *  - State lives in an opaque swr_system (device buffers owned by the
 *    library) instead of caller-owned host arrays inside packs.  Host views
 *    are copied in with swr_set_atoms() and out with swr_get(); swr_get()
 *    returns lists in the reference's record layouts (swreax_layout.h).
 *  - Every function returns an int status (SWR_OK or an SWR_ERR_* code)
 *    instead of void + MPI_Abort.  Capacity errors carry the required size
 *    (swr_get_scalar "need_bond_cap", ...), so a caller can grow and retry.
 *  - Calls enqueue on the system's HIP stream and return; swr_sync() joins
 *    (the analogue of athread_join) and reports device-side errors.
 *  - The CG loop of FixQEqReaxSunway::CG_v2 sits behind one call that
 *    solves s and t together on the device.
 *
 * No CUDA, no torch types: plain pointers and sizes.
 */
#ifndef SWREAX_H
#define SWREAX_H

#include <stdint.h>

#include "swreax_layout.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes */
enum {
  SWR_OK = 0,
  SWR_ERR_ARG = 1,         /* bad argument / call order                      */
  SWR_ERR_BOND_CAP = 2,    /* per-atom or total bond capacity exceeded        */
  SWR_ERR_HBOND_CAP = 3,   /* per-atom or total hbond capacity exceeded       */
  SWR_ERR_NOMEM = 4,       /* device allocation failed                        */
  SWR_ERR_FFIELD = 5,      /* force-field file unreadable or type unmapped    */
  SWR_ERR_STATE = 6,       /* stage called before its inputs exist            */
  SWR_ERR_BIN = 7,         /* atom outside the neighbour bin grid             */
  SWR_ERR_NUMERIC = 8,     /* non-finite CG scalar                            */
  SWR_ERR_HIP = 10         /* HIP runtime error (see swr_last_error)          */
};

/* Settings of `pair_style reax/c NULL [keywords]` plus `fix qeq/reax nevery
 * swa swb tol reax/c` (pair_reaxc_sunway.cpp:202-290,
 * fix_qeq_reax_sunway.cpp:76-99).  swr_control_default() fills the NULL
 * defaults: bond_cut 5, hbond_cut 7.5, safezone 1.2, saferzone 1.4, mincap
 * 50, swa 0, swb 10, tol 1e-6, skin 2. */
typedef struct swr_control_t {
  double bond_cut, hbond_cut, thb_cut;
  double safezone, saferzone;
  int mincap, maxfar;
  double swa, swb, tol;
  double skin;          /* neighbor skin; cutneigh = max(nonb,hbond,2*bond) + skin */
  int qeq_mixed;        /* 1: FP32 inner CG on an FP32 copy of H, FP64 refinement */
  int qeq_imax;         /* CG iteration cap, 200 in the reference              */
  int enobonds;         /* pair_style keyword enobonds (1 = yes, the default)  */
  double thb_cutsq;     /* valence-angle BO product cutoff, 0.00001 (:219)    */
} swr_control_t;

typedef struct swr_system swr_system;

void        swr_control_default(swr_control_t *ctl);
const char *swr_last_error(void);

/* Create a system: reads the force field (Read_Force_Field,
 * reaxc_ffield_sunway.cpp:35) and maps LAMMPS types 1..ntypes to the named
 * elements (PairReaxCSunway::coeff, pair_reaxc_sunway.cpp:294-362). */
int swr_create(const char *ffield_path, int ntypes, const char *const *elements,
               const swr_control_t *ctl, swr_system **out);
int swr_destroy(swr_system *sys);

/* Atom state from host arrays (LAMMPS atom->tag/type/x/q, nlocal + nghost).
 * owner[g] is the local atom that ghost g images (owner[i] = i locally) and
 * img[3*g+d] in {-1,0,1} its periodic shift; owner/img replace LAMMPS's
 * forward-comm maps (fix_qeq_reax_sunway.cpp:1759-1813). */
int swr_set_atoms(swr_system *sys, int n, int N, const int *tag, const int *type,
                  const double *x, const double *q, const int *owner,
                  const signed char *img);
/* The atoms of the last swr_set_atoms are a new configuration, not the next
 * MD step of the previous one: clears the QEq warm-start history (s, t and
 * their 5-step history, fix_qeq_reax_sunway.cpp:518-535 at first setup) and
 * re-estimates the list capacities at the next Reset (Estimate_Storages,
 * reaxc_forces_sunway.cpp:1190-1292).  A caller stepping one system keeps
 * both by not calling it. */
int swr_new_atoms(swr_system *sys);

/* Per-atom capacities (fix REAXC num_bonds/num_hbonds, fix_reaxc_sunway.cpp:39-161). */
int swr_set_capacities(swr_system *sys, const int *num_bonds, const int *num_hbonds);

/* a1: full neighbour list of local + ghost atoms on a GPU cell list.
 * Replaces npair_full_bin_ghost_sunway_build_packed (npair_full_bin_ghost_sw5.c:64;
 * CPE kernel :80-226).  cutneigh <= 0 uses the control's cutmax + skin. */
int swr_npair_full_bin_ghost_build(swr_system *sys, double cutneigh);

/* a1': the same list for local atoms only; ghost rows are empty (gnum = 0).
 * Replaces npair_full_bin_atomonly_sunway_build_packed
 * (npair_full_bin_atomonly_sw5.c:6, CPE kernel :57-180, called from
 * NPairFullBinAtomonlySunway::build, npair_full_bin_atomonly_sunway.cpp:151). */
int swr_npair_full_bin_atomonly_build(swr_system *sys, double cutneigh);

/* a2: pack tag, element type, x, q (write_reax_atoms_and_pack,
 * pair_reaxc_sw64.c:13; caller pair_reaxc_sunway.cpp:818). */
int swr_write_reax_atoms_and_pack(swr_system *sys);

/* a4: Reset — Hindex/numH, bond and hbond start/end from the capacities
 * (Reset, reaxc_reset_tools_sunway.cpp:200-211).  On the first call, if no
 * capacities were set, runs the setup-time estimate (Estimate_Storages,
 * reaxc_forces_sunway.cpp:1159; Init_Lists, reaxc_init_md_sunway.cpp:171). */
int swr_reset(swr_system *sys);

/* a3: far-neighbour list, d^2 <= nonb_cut^2 (write_reax_lists_c,
 * pair_reaxc_sw64.c:42; caller pair_reaxc_sunway.cpp:906). */
int swr_write_reax_lists(swr_system *sys);

/* a5+a6: bonds with uncorrected BO and dBO terms, sorted by nbr, plus
 * hbonds of local H, from the far list (Init_Forces_noQEq_Full_C,
 * reaxc_forces_sw64.c:594; caller reaxc_forces_sunway.cpp:845). */
int swr_init_forces_noqeq_full(swr_system *sys);

/* a3+a5+a6 in one pass over the Verlet list: identical outputs to
 * swr_write_reax_lists() followed by swr_init_forces_noqeq_full(), without
 * re-reading the far list. */
int swr_write_reax_lists_init_forces(swr_system *sys);

/* a7: per-atom capacities for the next step from this step's counts, and
 * the overflow check (Validate_Lists_C, reaxc_forces_sw64.c:866-930): an
 * end index past the next atom's start returns SWR_ERR_BOND_CAP /
 * SWR_ERR_HBOND_CAP (the reference aborts). */
int swr_validate_lists(swr_system *sys);

/* a7: bond_mark[i] = min(bond_mark[i], bond_mark[j] + 1) over i's bonds in
 * one pass in atom order (sort_bonds_C, reaxc_forces_sw64.c:932-971); the
 * result equals the reference's sequential pass exactly. */
int swr_sort_bonds(swr_system *sys);

/* a8-a10: corrected bond orders, dBO coefficients, Delta and lone-pair terms
 * (BO_C, reaxc_bond_orders_sw64.c:487; caller reaxc_bond_orders_sunway.cpp:454). */
int swr_bo(swr_system *sys);

/* a11-a13: H matrix (full rows of local atoms, taper x shielded Coulomb),
 * preconditioner, right-hand sides and extrapolated warm starts
 * (FixQEqReaxSunway::init_matvec, fix_qeq_reax_sunway.cpp:618-723, with
 * compute_H_Full_C, fix_qeq_reax_sw64.c:75).  jlist holds the owning local
 * atom of each column (ghost columns folded onto their owners). */
int swr_qeq_init_matvec(swr_system *sys);

/* a14: b[i] = eta_i x_i + sum_j H_ij x_j for local rows, device pointers of
 * length n (sparse_matvec_C, fix_qeq_reax_sw64.c:42; _spawn/_join :67,:71;
 * callers fix_qeq_reax_sunway.cpp:1068-1143).  Asynchronous like _spawn;
 * swr_sync() is the _join. */
int swr_sparse_matvec(swr_system *sys, const double *x_dev, double *b_dev);
/* The same with host arrays of length n (the reference's sparse_matvec takes
 * the fix's host vectors, fix_qeq_reax_sunway.cpp:1563). Synchronises. */
int swr_sparse_matvec_h(swr_system *sys, const double *x_host, double *b_host);

/* a15: both CG_v2 solves (s for b = -chi, t for b = -1) with one shared H
 * read per iteration (FixQEqReaxSunway::CG_v2, fix_qeq_reax_sunway.cpp:983;
 * called at :576-577).  Synchronises; writes the iteration counts. */
int swr_qeq_cg(swr_system *sys, int *iters_s, int *iters_t);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the FP64 solve of swr_qeq_cg on sys's H, allocating
 * nothing.  b (right-hand sides) and x (warm start in, solution out) are
 * device arrays of n (s, t) pairs, interleaved; every other buffer the
 * kernels use (CG vectors, partial dots, scalars, counts, layout) lives in a
 * per-GPU static device array of about 5.8 MB, held until the process ends.
 * At most 32768 local atoms, one system, FP64 QEq (SWR_ERR_ARG otherwise);
 * needs swr_qeq_init_matvec first.  Synchronises; writes the iteration
 * counts.  Every call on one GPU shares that array: calls that overlap on
 * one GPU race. */
int swr_qeq_cg_noalloc(swr_system *sys, const double *b, double *x, int *iters_s, int *iters_t);

/* a16: q = s - (sum s / sum t) t, history shift, ghost q refresh
 * (FixQEqReaxSunway::calculate_Q, fix_qeq_reax_sunway.cpp:1697-1755). */
int swr_calculate_q(swr_system *sys);

/* fix qeq/reax pre_force: init_matvec + CG(s,t) + calculate_Q
 * (FixQEqReaxSunway::pre_force, fix_qeq_reax_sunway.cpp:539-600).
 * iters may be NULL (then nothing synchronises). */
int swr_qeq_pre_force(swr_system *sys, int *iters_s, int *iters_t);

/* Pair compute stages in the reference order (PairReaxCSunway::compute,
 * pair_reaxc_sunway.cpp:541-793, through BO): pack, Reset, lists + bonds,
 * Validate_Lists, sort_bonds, BO. */
int swr_pair_compute(swr_system *sys);

/* Batch of independent systems in one swr_system (no reference counterpart:
 * the batched and serving entry points of this build).  The local atoms of
 * system k are rows [row_offsets[k], row_offsets[k+1]) (row_offsets[0] = 0,
 * row_offsets[nsys] = n); their ghosts may come in any order after n, with
 * owner[] pointing at the system's own local atoms.  Systems must be placed
 * so that no atom of one lies within the neighbour cutoff of another.  Every
 * stage then runs once over all systems (one launch per stage); QEq keeps
 * per-system CG scalars, stopping tests and charge neutrality, so each
 * system's charges are those it gets alone.  Per-system iteration counts:
 * swr_get(sys, "sys_iters", ...) (2 ints per system: s, t).  Reset to one
 * system whenever swr_set_atoms changes the local atom count.  The
 * mixed-precision solve keeps per-system refinement and inner-solve state
 * the same way. */
int swr_set_batch(swr_system *sys, int nsys, const int *row_offsets);

/* Device-resident QEq warm start of one system between evaluations (no
 * reference counterpart: fix qeq/reax keeps s_hist/t_hist in its own arrays
 * across the steps of one run, fix_qeq_reax_sunway.cpp:711-714; a serving
 * loop evaluates a resident system's successive steps in different batches).
 * swr_hist_save copies the s and t histories (5 per row each) of local rows
 * [row0, row0+n) into *store, a device buffer that outlives the batch
 * (allocated when *store is NULL or smaller than n rows, else reused);
 * fill != 0 sets all five slots of each row to slot 0 (the solution just
 * reached), so the extrapolation of the next solve starts from it.
 * swr_hist_load copies a store's n rows into rows [row0, row0+n) of sys
 * (after swr_set_atoms/swr_new_atoms, before the solve).  Both enqueue on
 * sys's stream only; a store is read by a load on another system's stream
 * only after the saving system was synchronised.  swr_hist_free releases a
 * store no call in flight uses. */
typedef struct swr_hist_store swr_hist_store;
int swr_hist_save(swr_system *sys, int row0, int n, int fill, swr_hist_store **store);
int swr_hist_load(swr_system *sys, int row0, int n, const swr_hist_store *store);
int swr_hist_free(swr_hist_store *store);

/* Non-bonded pass (vdW_Coulomb_Energy_Full_C, reaxc_nonbonded_sw64.c:260-500;
 * host form :522-830; the vdw_type 1 CPE kernel reaxc_nonbonded_cpe.h:1-420):
 * over the local atoms' far lists and the packed charges, the tapered
 * shielded Morse vdW (+ inner wall for vdw_type 2/3) and shielded Coulomb
 * energies.  Adds each local atom's gradient -(CEvd + CEclmb) dvec summed
 * over its far entries to fCdDelta[i][0..2] (zeroed by swr_reset, as
 * Reset_Workspace does), and leaves eng_vdwl, eng_coul and the virial (half
 * of every full-list pair) in swr_get(sys, "nonbonded_energy") (8 doubles),
 * and each local atom's half of its pairs' energy in "nonbonded_eatom".
 * Run after swr_write_reax_lists (or the fused list pass) of the step;
 * enqueues only.  lgvdw is not supported (the pair_style default is off). */
int swr_vdw_coulomb_energy(swr_system *sys);

/* Bond and atom energies: Atom_Energy (reaxc_multi_body_sunway.cpp:41-57),
 * i.e. Merge_Bonds_Atom_Energy_C_New (reaxc_multi_body_sw64.c:21-337; the
 * separate Bonds() returns at once, reaxc_bonds_sunway.cpp:49-52).  Over the
 * local atoms: polarization, lone-pair (+ C-C correction), over- and
 * under-coordination energies and, for each bond a local atom owns (smaller
 * orig_id, position order between images of one atom), the bond energy and
 * the terminal triple-bond stabilisation.  Adds to Cdbo/Cdbopi/Cdbopi2 of
 * the bond slots and to CdDelta (fCdDelta[i][3]) of the atoms and their
 * bonded neighbours (ghosts included); leaves e_bond, e_lp, e_ov, e_un,
 * e_pol in swr_get(sys, "bonded_energy") (5 doubles; e_pol is the plain sum
 * over atoms, the value the reference adds to eng_coul) and each local
 * atom's share in "bonded_eatom" (5 rows of be_stride doubles, the row
 * stride being swr_get_scalar(sys, "bonded_stride")).  Run after swr_bo of
 * the step; enqueues only.  The pair_style keyword enobonds is
 * swr_control_t.enobonds. */
int swr_atom_energy(swr_system *sys);

/* Hydrogen bonds: Hydrogen_Bonds (reaxc_forces_sunway.cpp:123-129) =
 * Hydrogen_Bonds_C (reaxc_hydrogen_bonds_sw64.c:26-283): for every local H
 * atom j, bond j-i to an acceptor with BO >= 0.01 and hydrogen-bond partner
 * k of j (orig_id(k) != orig_id(i), r0_hb(i,j,k) > 0 in the force field's
 * hydrogen-bond block): the energy, CEhb1 into Cdbo of j's slot, and the
 * angle and distance terms into fCdDelta[.][0..2] of i, j and k.  Leaves
 * e_hb in swr_get(sys, "hbond_energy") and each local H atom's share in
 * "hbond_eatom".  Run after swr_atom_energy and before
 * swr_add_dbond_forces, as the reference orders them; enqueues only. */
int swr_hydrogen_bonds(swr_system *sys);

/* Valence angles: the valence part of Merge_Torsion_Valence_Angles
 * (reaxc_torsion_angles_sw64.c:577-1117, host form :812-1117), which
 * Torsion_Angles calls (reaxc_torsion_angles_sunway.cpp:538-559;
 * Valence_Angles itself returns at once, reaxc_valence_angles_sunway.cpp:92-96).  For every
 * local atom j and pair of its bonds j-k, j-h with BO > thb_cut and
 * BO_jk BO_hj > thb_cutsq, and every parameter set of the force field's
 * valence-angle block for (k, j, h) with |p_val1| > 0.001: angle, penalty
 * and coalition energies, their coefficients into Cdbo/Cdbopi/Cdbopi2 of j's
 * bonds and CdDelta of j, k, h, and the angle derivative into
 * fCdDelta[.][0..2] of j, k, h.  Leaves e_ang, e_pen, e_coa in
 * swr_get(sys, "valence_energy") and each local atom's share in
 * "valence_eatom" (3 rows of n doubles).  The torsion part of the same
 * function is swr_torsion_angles.  Run after swr_bo and before
 * swr_add_dbond_forces; enqueues only. */
int swr_valence_angles(swr_system *sys);

/* Torsions: the torsion part of Merge_Torsion_Valence_Angles
 * (reaxc_torsion_angles_sw64.c:812-1408; Calculate_Omega :21-104).  For every
 * local atom j and bond j-k (BO > thb_cut), once per physical bond (the
 * smaller orig_id first, images of one atom by position), every other bond
 * j-i and bond k-l (l != i) with BO > thb_cut, BO_ij BO_jk BO_kl > thb_cut
 * and a torsion entry for (i, j, k, l) in the force field's torsion block:
 * torsion and four-body conjugation energies, their coefficients into
 * Cdbopi of j-k, Cdbo of j-i, j-k, k-l and CdDelta of j and k, and the angle
 * and dihedral derivatives into fCdDelta[.][0..2] of i, j, k, l.  Leaves
 * e_tor, e_con in swr_get(sys, "torsion_energy") and each local atom's share
 * in "torsion_eatom" (2 rows of n doubles).  Run after swr_bo and before
 * swr_add_dbond_forces; enqueues only. */
int swr_torsion_angles(swr_system *sys);

/* Add_All_dBond_to_Forces_C (reaxc_forces_sw64.c:417-589): over all N atoms
 * and their bonds, the chain rule of the bond-order derivatives with the
 * coefficients Cdbo/Cdbopi/Cdbopi2 of both halves of each bond and the
 * CdDelta of both atoms, added to fCdDelta[.][0..2] (the energy gradient;
 * ghost entries are to be summed onto their owners, as LAMMPS' reverse
 * communication does).  Run after the terms that produce those
 * coefficients (swr_atom_energy); enqueues only. */
int swr_add_dbond_forces(swr_system *sys);

/* Forces on the local atoms: what PairReaxCSunway::compute does with the
 * gradient (pair_reaxc_sunway.cpp:746-760, f[k] += -fCdDelta[k] over all N
 * atoms) followed by LAMMPS' reverse communication, which adds each ghost's
 * force onto the local atom it images (owner[] of swr_set_atoms, or the
 * harness's borders).  f[i] = -(fCdDelta[i] + sum over the ghosts g of i of
 * fCdDelta[g]), summed in ascending ghost order; read back with
 * swr_get(sys, "f") ([n][3]).  Run after the terms that fill fCdDelta
 * (swr_vdw_coulomb_energy, swr_hydrogen_bonds, swr_add_dbond_forces);
 * enqueues only. */
int swr_compute_forces(swr_system *sys);

/* atom->mass: mass[1..ntypes] per atom type (mass[0] unused, as in LAMMPS).
 * Defaults to the force field's mass of each type's element. */
int swr_set_masses(swr_system *sys, int ntypes, const double *mass);

/* FixNVESunway::initial_integrate / final_integrate =
 * fix_nve_initial_integrate_para / fix_nve_final_integrate_para
 * (fix_nve_sw64.c:25-173; per-type mass branch): for every local atom
 * v += dtf/m f, x += dtv v (initial) and v += dtf/m f (final), with f from
 * the last swr_compute_forces and v from swr_put(sys, "v", ...).  As in the
 * reference, dtv = dt and dtf = dt/2 * ftm2v (fix_nve_sunway.cpp:56-57;
 * ftm2v = 1/48.88821291^2 in real units).  Ghost positions are refreshed by
 * the caller's forward communication or rebuild, as in LAMMPS.  Fails with
 * "All masses are not set" when a type has no positive mass; enqueues only. */
int swr_fix_nve_initial_integrate(swr_system *sys, double dtv, double dtf);
int swr_fix_nve_final_integrate(swr_system *sys, double dtv, double dtf);

/* Host-side packing of nsys independent systems into one batch (no
 * reference counterpart; host code, no device calls).  System k has nat[k]
 * local atoms, listed in order in x (3 per atom), type (1-based, already in
 * the batch's element numbering) and tag; its box is lo/hi[3k..3k+2].  A
 * periodic system gets ghost images within cutghost in LAMMPS
 * Comm::borders order (per dimension, atoms with x <= lo+cut imaged at +L,
 * then x >= hi-cut at -L, over the atoms present when the dimension starts).
 * Each system is then translated into its own cell of an m x m x m grid
 * (m = ceil(nsys^(1/3)), cell edge = the largest extent + cutghost + 1), so
 * no atom lies within cutghost of another system's.  Outputs: local atoms of
 * all systems first (rows row_off[k]..row_off[k+1]), then their ghosts
 * (n + ghost_off[k]...); tag of system k offset by row_off[k]; owner = the
 * local row a ghost copies; img = its periodic image; shift[3k..] = the
 * translation applied.  *N_out = n + ghosts.  With x_out == NULL only
 * *N_out is computed; otherwise cap must be >= that N. */
int swr_batch_pack(int nsys, const int *nat, const double *x, const int *type, const int *tag,
                   const double *lo, const double *hi, const int *periodic, double cutghost, long long cap,
                   double *x_out, int *type_out, int *tag_out, int *owner_out, signed char *img_out,
                   int *row_off, int *ghost_off, double *shift, long long *N_out);

/* ---- autotuner of launch shapes (no reference counterpart: the Sunway
 * kernels have fixed ISTEP/SNSTEP tiles).  Families: "spmv_fp64" (the FP64
 * CG's SpMV), "spmv_fp32" (the mixed-precision inner SpMV), "spmv_resid"
 * (its FP64 residual) with shapes T512G8U4 T512G8U2 T256G8U4 T256G4U4
 * (threads per workgroup, lanes per H row, entries in flight per lane), and
 * "far" (the far pass) with G4W2 G4W4 G6W2 G8W4 (gatherer and writer waves
 * per workgroup).
 * "panels" (on, off: the SpMV's LDS column lists) is set with swr_set_shape
 * only.  swr_autotune times every shape the system's current state allows
 * (reps launches each, HIP events) and keeps the fastest; run it between
 * steps.  A shape changes the order of sums inside an H row, never a list.
 * With the environment variable SWR_TUNE_CACHE naming a file, the choices
 * are appended there (device, family, log2 of the local atom count) and
 * swr_autotune_load applies those matching the system (returns how many). */
int     swr_autotune(swr_system *sys, int reps);
int     swr_autotune_load(swr_system *sys);
int     swr_set_shape(swr_system *sys, const char *family, const char *shape);
int     swr_get_shape(swr_system *sys, const char *family, char *out, int len);
/* JSON of the last swr_autotune (microseconds per shape, the choice);
 * returns its length, copies at most len - 1 bytes into out. */
int64_t swr_tune_report(swr_system *sys, char *out, int64_t len);

/* Join the stream; returns the first device-raised error. */
int swr_sync(swr_system *sys);
/* hipDeviceSynchronize on the current device (a barrier for callers that
 * time several systems or streams). */
int swr_device_synchronize(void);

/* Mirrors: download a named device array into host memory in the
 * reference's layout; returns bytes (or -1 for an unknown name).  Pass
 * out = NULL to query the size.  Names: x q tag type owner packed_atoms
 * verlet_offsets verlet_list far far_index far_end num_bonds num_hbonds
 * Hindex bond_index bond_end bonds bo_data BO BOpi Cdbo Cdbopi Cdbopi2
 * hbond_index hbond_end hbonds hbond_far bo_dboc dDeltap_self bond_mark
 * Deltap Deltap_boc Delta Delta_e Delta_val vlpex nlp nlp_temp Clp Delta_lp
 * dDelta_lp Delta_lp_temp dDelta_lp_temp H_firstnbr H_numnbrs H_jlist H_val
 * Hdia_inv b_s b_t s t s_hist t_hist fCdDelta nonbonded_energy nonbonded_eatom bonded_energy
 * bonded_eatom hbond_energy hbond_eatom.  hbonds' ptr
 * fields come back as
 * device addresses; hbond_far gives each hbond's far-list slot. */
int64_t swr_get(swr_system *sys, const char *name, void *out);
int     swr_put(swr_system *sys, const char *name, const void *in, int64_t bytes);
int64_t swr_get_scalar(swr_system *sys, const char *name);
/* Device pointer of a named array (for callers that keep data resident). */
void   *swr_device_ptr(swr_system *sys, const char *name);

/* Per-kernel timing with HIP events on the system's stream.  Slots: verlet
 * bins far bonds reset validate mark bo build_H cg_setup cg_iter cg_update
 * cg_scalars spmv qeq_q mixed_inner mixed_outer.  read() synchronises, returns the accumulated
 * milliseconds and launch count, and clears them when reset != 0. */
int swr_ktimer_enable(swr_system *sys, int on);
int swr_ktimer_read(swr_system *sys, const char *slot, double *ms, long long *count, int reset);

/* ---- mini-LAMMPS host harness (C++ classes behind a C handle) ----------
 * Supplies what LAMMPS would around the path: box, ghost images (borders /
 * forward comm on the device), the reneighbour trigger and NVE.  Commands
 * follow LAMMPS syntax for the subset this path needs:
 *   pair_style reax/c/sunway NULL [checkqeq yes|no] [safezone v] [mincap n] [maxfar n]
 *   pair_coeff * * <ffield> <element per type>
 *   fix <id> all qeq/reax/sunway <nevery> <swa> <swb> <tol> reax/c [mixed]
 *   neighbor <skin> bin
 *   neigh_modify every <k>
 *   timestep <dt_fs> */
typedef struct swr_lmp swr_lmp;
int         swr_lmp_open(swr_lmp **out);
int         swr_lmp_close(swr_lmp *lmp);
int         swr_lmp_command(swr_lmp *lmp, const char *cmd);
int         swr_lmp_create_atoms(swr_lmp *lmp, int n, const double *lo, const double *hi,
                                 int periodic, const double *x, const double *v,
                                 const int *type, const int *tag);
/* Optional, before swr_lmp_setup: QEq s/t histories (n x 5 each, atom-major,
 * the layout of fix qeq/reax's s_hist/t_hist) to warm-start the setup solve,
 * as a restart file's fix state would. */
int         swr_lmp_set_history(swr_lmp *lmp, int n, const double *s_hist, const double *t_hist);
int         swr_lmp_setup(swr_lmp *lmp);
int         swr_lmp_run(swr_lmp *lmp, int nsteps);
swr_system *swr_lmp_system(swr_lmp *lmp);
double      swr_lmp_stat(swr_lmp *lmp, const char *name);

#ifdef __cplusplus
}
#endif

#endif /* SWREAX_H */
