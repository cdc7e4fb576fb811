/*
 * swreax_layout.h — byte layouts shared by the MI355X path, its C-ABI callers
 * and the CPU oracle.
 *
 * Every struct here is byte-compatible with the reference's type of the same
 * role, so a LAMMPS-side caller can hand the reference's own arrays across
 * (or memcpy a device mirror into them) without repacking:
 *
 *   swr_atom_pack_t          <-> atom_pack_t              reaxc_ctypes_sunway.h:233-238   (40 B)
 *   swr_far_neighbor_t       <-> far_neighbor_data_full   reaxc_ctypes_sunway.h:575-583   (64 B)
 *   swr_hbond_t              <-> hbond_data               reaxc_ctypes_sunway.h:584-588   (16 B)
 *   swr_bond_order_t         <-> bond_order_data          reaxc_ctypes_sunway.h:612-624   (200 B)
 *   swr_bond_t               <-> bond_data                reaxc_ctypes_sunway.h:643-651   (56 B)
 *   swr_sparse_matrix_t      <-> FixQEqReaxSunway::sparse_matrix  fix_qeq_reax_sunway.h:87-94
 *
 * Names carry an swr_ prefix so that this header can be included next to the
 * reference's headers in one translation unit.
 */
#ifndef SWREAX_LAYOUT_H
#define SWREAX_LAYOUT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef double swr_rvec[3];
typedef double swr_rvec2[2];
typedef int    swr_ivec[3];

/* Packed per-atom record written each step before the list stages
 * (write_reax_atoms_and_pack_para, pair_reaxc_sw64.c:114-190). */
typedef struct swr_atom_pack_t {
  int      orig_id;   /* LAMMPS tag (rc_tagint = int under LAMMPS_SMALLBIG) */
  int      type;      /* force-field element index, map[LAMMPS type]      */
  swr_rvec x;
  double   q;
} swr_atom_pack_t;

/* One far-neighbour entry (d^2 <= nonb_cut^2), written by the far-list stage
 * (write_reax_lists_c_para, pair_reaxc_sw64.c:286-310). */
typedef struct swr_far_neighbor_t {
  int      nbr;
  int      type;
  int      orig_id;
  swr_ivec rel_box;   /* always 0: ghosts are explicit images */
  double   d;
  double   q;
  swr_rvec dvec;      /* x_j - x_i */
} swr_far_neighbor_t;

/* Hydrogen-bond entry: ptr points at the far-list entry of the pair
 * (Init_Forces_noQEq_HB_Full_C, reaxc_forces_sw64.c:850-853).  Device lists
 * hold device addresses; host mirrors are rebased onto the host far list. */
typedef struct swr_hbond_t {
  int                 nbr;
  int                 scl;
  swr_far_neighbor_t *ptr;
} swr_hbond_t;

/* Per-bond derivative block (reaxc_ctypes_sunway.h:612-624). */
typedef struct swr_bond_order_t {
  double   BO_s;
  swr_rvec dBOp, dln_BOp_s, dln_BOp_pi, dln_BOp_pi2;
  double   C1dbo, C2dbo, C3dbo;
  double   C1dbopi, C2dbopi, C3dbopi, C4dbopi;
  double   C1dbopi2, C2dbopi2, C3dbopi2, C4dbopi2;
  double  *CdboReduction;   /* unused on this path, kept for layout */
} swr_bond_order_t;

/* Bond entry (reaxc_ctypes_sunway.h:643-651). */
typedef struct swr_bond_t {
  double   d;
  swr_rvec dvec;
  int      nbr;
  int      dbond_index;   /* slot the bond was appended at, before the per-atom sort */
  swr_ivec rel_box;
  int      sym_index;     /* slot of the reverse bond j->i in j's list */
} swr_bond_t;

/* QEq matrix (fix_qeq_reax_sunway.h:87-94): row i occupies
 * jlist/val[firstnbr[i] .. firstnbr[i]+numnbrs[i]). */
typedef struct swr_sparse_matrix_t {
  int    *firstnbr;
  int    *numnbrs;
  int    *jlist;
  double *val;
} swr_sparse_matrix_t;

#ifdef __cplusplus
}  /* extern "C" */

static_assert(sizeof(swr_atom_pack_t) == 40, "atom_pack_t is 40 B");
static_assert(offsetof(swr_atom_pack_t, x) == 8 && offsetof(swr_atom_pack_t, q) == 32, "atom_pack_t offsets");
static_assert(sizeof(swr_far_neighbor_t) == 64, "far_neighbor_data_full is 64 B");
static_assert(offsetof(swr_far_neighbor_t, rel_box) == 12 && offsetof(swr_far_neighbor_t, d) == 24 &&
              offsetof(swr_far_neighbor_t, q) == 32 && offsetof(swr_far_neighbor_t, dvec) == 40,
              "far_neighbor_data_full offsets");
static_assert(sizeof(swr_hbond_t) == 16 && offsetof(swr_hbond_t, ptr) == 8, "hbond_data is 16 B");
static_assert(sizeof(swr_bond_order_t) == 200, "bond_order_data is 200 B");
static_assert(offsetof(swr_bond_order_t, dBOp) == 8 && offsetof(swr_bond_order_t, C1dbo) == 104 &&
              offsetof(swr_bond_order_t, C1dbopi2) == 160 && offsetof(swr_bond_order_t, CdboReduction) == 192,
              "bond_order_data offsets");
static_assert(sizeof(swr_bond_t) == 56, "bond_data is 56 B");
static_assert(offsetof(swr_bond_t, nbr) == 32 && offsetof(swr_bond_t, dbond_index) == 36 &&
              offsetof(swr_bond_t, rel_box) == 40 && offsetof(swr_bond_t, sym_index) == 52,
              "bond_data offsets");
#endif

#endif /* SWREAX_LAYOUT_H */
