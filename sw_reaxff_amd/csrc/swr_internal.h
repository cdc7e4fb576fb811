// swr_internal.h — device-side state and helpers of the MI355X path.
//
// One swr_system owns every device buffer of one simulated system: atoms
// (local + ghost images), the Verlet list, the far/bond/hbond lists in the
// reference's record layouts, the BO workspace and the QEq matrix and solver
// vectors.  All kernels run on the system's stream; entry points enqueue and
// return, and swr_sync() reports errors raised on the device.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/swreax.h"
#include "swr_ffield.h"

#define SWR_MAX_ELEM 8    // force-field elements usable per system
#define SWR_MAX_TYPES 8   // LAMMPS atom types per system
#define SWR_PANEL_ROWS 64      // H rows per SpMV panel
#define SWR_PANEL_MAXC 3072    // distinct columns a panel's LDS list holds
#define SWR_WAVE 64

// ---- error flags raised by kernels (bit mask in DevStatus::flags)
#define SWR_F_BOND_CAP   (1u << 0)   // an atom has more bonds than num_bonds[i]
#define SWR_F_HBOND_CAP  (1u << 1)   // an H has more hbonds than num_hbonds[i]
#define SWR_F_BOND_TOTAL (1u << 2)   // sum of bond capacities exceeds the list
#define SWR_F_HBOND_TOTAL (1u << 3)  // sum of hbond capacities exceeds the list
#define SWR_F_BOND_LDS   (1u << 4)   // more than SWR_MAX_ATOM_BONDS bonds for one atom
#define SWR_F_MARK_DEPTH (1u << 5)   // bond_mark relaxation did not converge
#define SWR_F_BIN        (1u << 6)   // an atom fell outside the bin grid
#define SWR_F_NAN        (1u << 7)   // non-finite value in CG scalars
#define SWR_F_FIXED_RANGE (1u << 9)  // a bonded term beyond the fixed-point accumulators' range
#define SWR_F_CG_MAXIT   (1u << 8)   // a QEq solve stopped at its iteration limit (a warning, as in the reference)

// far-pass gatherer/writer wave splits (lists.hip) the autotuner chooses from
enum { SWR_FAR_G4W2 = 0, SWR_FAR_G4W4, SWR_FAR_G6W2, SWR_FAR_G8W4, SWR_FAR_NSHAPES };

#define SWR_MAX_ATOM_BONDS 32        // bonds one atom may hold while building its list (the reference: MAX_LEN 20, MAXBOND 35)

#define HIP_CHECK(expr)                                                    \
  do {                                                                     \
    hipError_t _e = (expr);                                                \
    if (_e != hipSuccess) {                                                \
      swr_set_error_msg(std::string(#expr) + ": " + hipGetErrorString(_e)); \
      return SWR_ERR_HIP;                                                  \
    }                                                                      \
  } while (0)

void swr_set_error_msg(const std::string &m);

// Force-field parameters the path reads, element-indexed (sbp/tbp subset,
// reaxc_ctypes_sunway.h:100-170) plus QEq parameters per LAMMPS type.
struct DevTbp {
  double p_bo1, p_bo2, p_bo3, p_bo4, p_bo5, p_bo6;
  double r_s, r_p, r_pp;
  double p_boc3, p_boc4, p_boc5;
  double v13cor, ovc;
};

struct DevNb {
  double D, alpha, r_vdWinv, alpha_div_rvdW, powgi_vdW1;  // shielded Morse
  double gamma;                                           // Coulomb shielding
  double ecore, acore, rcoreinv, a_div_rcore;             // inner wall
};

// Hydrogen-bond parameters (acceptor i, hydrogen j, acceptor k); r0_hb <= 0: none.
struct DevHb {
  double r0_hb, p_hb1, p_hb2, p_hb3;
};

// Valence-angle parameter sets of one (end, centre, end) element triple,
// only those with |p_val1| > 0.001 (the ones the energy uses), in file order.
constexpr int SWR_MAX_THB = 5;

// Torsion parameters of one (i, j, k, l) element quadruple; cnt 0: no term.
struct DevTor {
  int cnt;
  double V1, V2, V3, p_tor1, p_cot1;
};
struct DevThb {
  double theta_00, p_val1, p_val2, p_coa1, p_val7, p_pen1, p_val4;
};

// Bond and atom energies (Merge_Bonds_Atom_Energy_C_New), element-pair indexed.
struct DevBe {
  double De_s, De_p, De_pp, p_be1, p_be2, p_ovun1;
};

struct DevParams {
  int nelem, ntypes;
  int map[SWR_MAX_TYPES + 1];
  double type_mass[SWR_MAX_TYPES + 1];   // atom->mass per LAMMPS type (fix nve)
  double r_s[SWR_MAX_ELEM], r_pi[SWR_MAX_ELEM], r_pi_pi[SWR_MAX_ELEM];
  double valency[SWR_MAX_ELEM], valency_e[SWR_MAX_ELEM], valency_boc[SWR_MAX_ELEM];
  double valency_val[SWR_MAX_ELEM], nlp_opt[SWR_MAX_ELEM], mass[SWR_MAX_ELEM];
  int p_hbond[SWR_MAX_ELEM];
  DevTbp tbp[SWR_MAX_ELEM * SWR_MAX_ELEM];
  // Per element pair, a distance beyond which BO' < bo_cut for certain (BO'
  // falls monotonically with d for these parameters); bond candidates are
  // pre-filtered by it and then tested exactly.  bond_cut when no bound holds.
  double rbo[SWR_MAX_ELEM * SWR_MAX_ELEM];
  double p_boc1, p_boc2, p_lp1;
  double bo_cut, bond_cut, hbond_cut, nonb_cut, nonb_cutsq;
  // QEq, LAMMPS-type indexed
  double chi[SWR_MAX_TYPES + 1], eta[SWR_MAX_TYPES + 1], gamma[SWR_MAX_TYPES + 1];
  double shld[(SWR_MAX_TYPES + 1) * (SWR_MAX_TYPES + 1)];
  double Tap[8];
  double swa, swb, swbsq;
  double tol;
  // Non-bonded (vdW_Coulomb_Energy_Full_C), element-pair indexed: the
  // factors its host form precomputes per pair (reaxc_nonbonded_sw64.c:
  // 386-402), the Coulomb shielding gamma, and the taper on
  // [nonb_low, nonb_cut] (Init_Taper, reaxc_init_md_sunway.cpp:100-134).
  DevNb nb[SWR_MAX_ELEM * SWR_MAX_ELEM];
  double nb_tap[8];
  double p_vdW1, p_vdW1i;
  int vdw_type;
  // Bond and atom energies: element parameters, the pair block, and the
  // general parameters l[3..10], l[31], l[32], l[37] they read
  double p_lp2[SWR_MAX_ELEM], p_ovun2[SWR_MAX_ELEM], p_ovun5[SWR_MAX_ELEM], chi_e[SWR_MAX_ELEM],
      eta_e[SWR_MAX_ELEM];
  int is_carbon[SWR_MAX_ELEM];
  DevBe be[SWR_MAX_ELEM * SWR_MAX_ELEM];
  double p_lp3, p_ovun3, p_ovun4, p_ovun6, p_ovun7, p_ovun8, gp3, gp4, gp7, gp10;
  int gp37, enobonds;
  DevHb hb[SWR_MAX_ELEM * SWR_MAX_ELEM * SWR_MAX_ELEM];
  // Valence angles: element p_val3/p_val5, general l[2,14,16,17,19,20,21,
  // 30,33,38], control thb_cut/thb_cutsq
  double p_val3[SWR_MAX_ELEM], p_val5[SWR_MAX_ELEM];
  double p_val6, p_val8, p_val9, p_val10, p_pen2, p_pen3, p_pen4, p_coa2, p_coa3, p_coa4, thb_cut, thb_cutsq;
  int thb_n[SWR_MAX_ELEM * SWR_MAX_ELEM * SWR_MAX_ELEM];
  int thb_pc[SWR_MAX_ELEM * SWR_MAX_ELEM * SWR_MAX_ELEM];   // 1: a set has p_pen1 or p_coa1 != 0
  DevThb thb[SWR_MAX_ELEM * SWR_MAX_ELEM * SWR_MAX_ELEM][SWR_MAX_THB];
  // Torsions: general l[23,24,25,27] and the quadruple parameters
  double p_tor2, p_tor3, p_tor4, p_cot2;
  DevTor tor[SWR_MAX_ELEM * SWR_MAX_ELEM * SWR_MAX_ELEM * SWR_MAX_ELEM];
};

// Small block of device-written results the host reads back asynchronously.
struct DevStatus {
  unsigned int flags;
  int need_bond_cap;     // largest per-atom bond count that overflowed
  int need_hbond_cap;
  int total_bonds;       // sum of num_bonds (Reset)
  int total_hbonds;
  int numH;
  int mark_changed;
  int pad0;
  long long num_far;     // far entries written (sum over atoms)
  long long H_nnz;
  int cg_iters[2];       // s, t
  int inner_iters[2];    // last FP32 inner solve of the mixed-precision refinement
  int next_total_bonds;  // sums of the capacities Validate_Lists wrote (next step's Reset totals)
  int next_total_hbonds;
};

// Pipelined-CG scalars of one right-hand side (CG_v2,
// fix_qeq_reax_sunway.cpp:1080-1152).  Written by the last workgroup of each
// iteration kernel, read by the next.
struct CgScalars {
  double alpha_x;     // alpha used in x += alpha p (value before the update)
  double alpha_a;     // alpha used in r -= alpha ss, ... (value after the update)
  double beta;
  double sig_old, heta;
  double dg0, dg1;    // latest reduced u.r, u.w
  double b_norm;
  double bdot, udot_r, udot_w;   // setup dots
  double tol;         // stopping tolerance (negative: nothing to solve)
  int active_b;       // B/C of the current iteration run
  int active_a;       // A of the next iteration runs
  int iter;           // reference loop counter i
  int done;
};

// Kernel timer slots.
enum { KT_VERLET = 0, KT_FAR, KT_BUILD_H, KT_CG_ITER, KT_CG_SETUP, KT_BO, KT_VALIDATE, KT_SPMV, KT_RESET,
       KT_QEQ_Q, KT_BINS, KT_MIXED_INNER, KT_MIXED_OUTER, KT_BONDS, KT_CG_SCALARS, KT_MARK, KT_CG_UPDATE, KT_NONBONDED, KT_ATOM_ENERGY,
       KT_DBOND, KT_HBONDS, KT_VALENCE, KT_TORSION, KT_PANEL, KT_COUNT };

struct swr_system {
  // ---- host-side configuration
  int device = 0;
  hipStream_t stream = nullptr;
  swr_ffield_t ff;               // parsed force field (host)
  swr_control_t ctl;
  DevParams hp;                  // host copy of device params
  DevParams *dp = nullptr;       // device params

  // ---- atoms (device)
  int n = 0, N = 0, ncap = 0, Ncap = 0;
  double *x = nullptr;           // [N][3]
  double *v = nullptr;           // [n][3] (harness)
  double *q = nullptr;           // [N]
  int *type = nullptr, *tag = nullptr, *owner = nullptr;  // [N]
  signed char *img = nullptr;    // [N][3] image shift of ghosts, in box lengths
  int *mask = nullptr;           // unused on this path (group all)
  swr_atom_pack_t *packed = nullptr;

  // ---- bins + Verlet list
  int nbin[3] = {0, 0, 0};
  double binlo[3] = {0, 0, 0}, binsize[3] = {0, 0, 0}, bininv[3] = {0, 0, 0};
  int nbins = 0, binscap = 0;
  size_t cap_stencil = 0, cap_bin_of = 0, cap_bp_id = 0, cap_bp_pos = 0, cap_vcount = 0, cap_voff = 0;
  int *bin_of = nullptr, *bin_count = nullptr, *bin_start = nullptr;
  int *bin_occ = nullptr;        // [0]: number of occupied bins, [1..]: their ids
  // pinned host buffers the done checks copy into (device-to-host copies
  // into pageable memory are staged through a bounce buffer)
  int *pin_done = nullptr;        // the CG's count of finished systems
  void *pin_mix = nullptr;        // the mixed solve's counters and state (SWR_PIN_MIX bytes)
  int *pin_iters = nullptr;       // iteration counts, 2 per system
  int pin_iters_cap = 0;
  size_t cap_bin_occ = 0;
  // counting sort by bin (radix passes staged in LDS): key/value ping-pong
  // buffers and the per-(digit, block) counts and starts
  int *rs_keys = nullptr, *rs_vals = nullptr, *rs_hist = nullptr;
  size_t cap_rs_keys = 0, cap_rs_vals = 0, cap_rs_hist = 0;
  double4 *bp_pos = nullptr;     // atoms sorted by bin: x, y, z, unused
  int *bp_id = nullptr;
  int nrows = 0;
  int4 *stencil_rows = nullptr;  // (dy, dz, sx, 0)
  double cutneigh = 12.0;
  long long *voff = nullptr;     // [N+1]
  int *vcount = nullptr;         // [N]
  int *vlist = nullptr;
  long long vcap = 0, vtotal = 0;

  // ---- far list (reax_list FAR_NBRS_FULL)
  swr_far_neighbor_t *far = nullptr;   // capacity vcap, atom i slots [voff[i], voff[i+1])
  int *far_index = nullptr, *far_end = nullptr;

  // ---- capacities, bonds, hbonds
  int *num_bonds = nullptr, *num_hbonds = nullptr, *Hindex = nullptr;
  int *bond_index = nullptr, *bond_end = nullptr;
  int bond_cap = 0;
  swr_bond_t *bonds = nullptr;
  int *bond_nbr = nullptr;         // [bond_cap] bonds[].nbr as a compact array (sym_index searches, marks)
  swr_bond_order_t *bo_data = nullptr;
  double *BO = nullptr, *BOpi = nullptr, *Cdbo = nullptr, *Cdbopi = nullptr, *Cdbopi2 = nullptr;
  int *hbond_index = nullptr, *hbond_end = nullptr;   // [ncap] by Hindex
  int hbond_cap = 0;
  swr_hbond_t *hbonds = nullptr;
  bool capacities_set = false;
  bool next_totals_known = false;  // status_host holds the totals of the current capacities
  bool status_copy_queued = false; // a status copy was queued after Validate_Lists (done at the next sync)
  // bond_mark relaxation whose convergence is checked at the next sync
  int *mark_chg = nullptr;         // [MARK_MAX_LAUNCH] device: launch k changed a mark
  int *mark_chg_host = nullptr;    // pinned copy
  int mark_launched = 0;           // launches of the pending relaxation (0: none pending)
  long long vtotal_local = -1;     // Verlet entries of the local rows (voff[n]), known at each build

  // ---- BO workspace
  double *fcd = nullptr;           // fCdDelta [N][4]: gradient xyz + CdDelta (zeroed by Reset)
  double *nb_partials = nullptr;   // non-bonded per-block energy/virial partials
  size_t nb_partials_cap = 0;
  double *nb_result = nullptr;     // eng_vdwl, eng_coul, virial[6] of the last non-bonded pass
  double *nb_eatom = nullptr;      // per local atom: its half of its pairs' vdW + Coulomb energy
  size_t nb_eatom_cap = 0;
  // bond and atom energies, dBO forces
  double *be_part = nullptr;       // [5][cap] per local atom energies
  double *be_ce = nullptr;         // [2][cap] CEover4, CEunder4 per local atom
  double *be_bufc = nullptr;       // [cap] per atom buf_c of the dBO chain rule
  double *be_result = nullptr;     // e_bond, e_lp, e_ov, e_un, e_pol of the last Atom_Energy
  size_t cap_be_part = 0, cap_be_ce = 0, cap_be_bufc = 0, cap_be_result = 0;
  double *hb_part = nullptr;        // [n] per local H atom hydrogen-bond energy
  double *hb_result = nullptr;     // e_hb of the last Hydrogen_Bonds
  unsigned long long *hb_acc = nullptr;   // [N][3] fixed-point hydrogen-bond terms (zero between calls)
  size_t cap_hb_part = 0, cap_hb_result = 0, cap_hb_acc = 0;
  // valence angles: per-atom energies [3][va_stride], totals, fixed-point
  // terms of the end atoms [N][4] (zero between calls)
  double *va_part = nullptr, *va_result = nullptr;
  unsigned long long *va_acc = nullptr;
  size_t cap_va_part = 0, cap_va_result = 0, cap_va_acc = 0, va_stride = 0;
  // the valence work list: per local atom its angle count (then its rank in
  // its count bucket), the centres ordered by count, and the counters
  int *va_cnt = nullptr, *va_order = nullptr, *va_ctl = nullptr;
  size_t cap_va_cnt = 0, cap_va_order = 0, cap_va_ctl = 0;
  // torsions: per-atom energies [2][to_stride], totals, fixed-point terms of
  // atoms [N][4] and of bond slots [bond_cap] (zero between calls)
  double *to_part = nullptr, *to_result = nullptr;
  unsigned long long *to_acc = nullptr, *to_sacc = nullptr;
  size_t cap_to_part = 0, cap_to_result = 0, cap_to_acc = 0, cap_to_sacc = 0, to_stride = 0;
  // the torsion work list: per bond slot its quadruple count (then its rank
  // in its count bucket), the central bonds (j, slot) ordered by count, per
  // slot the six row terms of its central bond, and the counters
  int *to_tc = nullptr, *to_task = nullptr, *to_ctl = nullptr;
  double *to_row = nullptr;
  size_t cap_to_tc = 0, cap_to_task = 0, cap_to_ctl = 0, cap_to_row = 0;
  // forces on local atoms (ghost rows of fCdDelta summed onto owners) and the
  // ghost-per-owner CSR that sums them in a fixed order
  double *f = nullptr;                     // [n][3]
  size_t cap_f = 0;
  bool have_f = false;
  int *gcsr_cnt = nullptr, *gcsr_start = nullptr, *gcsr_list = nullptr;
  size_t cap_gcsr_cnt = 0, cap_gcsr_start = 0, cap_gcsr_list = 0;
  bool gcsr_valid = false;
  int gcsr_n = 0, gcsr_N = 0;
  size_t be_stride = 0;            // row stride of be_part/be_ce (the local atom count of the last pass)
  double *bo_dboc = nullptr;       // [N][2]
  double *dDeltap_self = nullptr;  // [N][3]
  int *bond_mark = nullptr, *bond_mark2 = nullptr;
  int *counts = nullptr;           // [2][Ncap]: bonds found per atom, hbonds found per H
  int mark_depth_hint = 0;         // bond_mark sweeps needed last time
  int *cand = nullptr;             // [Ncap][64]: far slots of bond candidates
  int *cand_count = nullptr;
  size_t cand_cap = 0;
  double *ws = nullptr;            // 13 arrays of N doubles, see WS_* below
  // ---- QEq
  // SpMV panels (panel.hip): the local rows in Morton order of their bins,
  // grouped by system, cut into panels of at most SWR_PANEL_ROWS rows of
  // one system; a panel's distinct H columns (at most SWR_PANEL_MAXC) are
  // listed in pcol, and each H entry holds its column's index in that list
  // in jloc, so the SpMV gathers x from LDS.  pcnt[p] < 0: panel p gathers
  // from global memory (more columns than the list holds, or panels off).
  int *prow = nullptr, *pbeg = nullptr, *pcnt = nullptr, *pcol = nullptr, *pkey = nullptr;
  int *sys_panel = nullptr;   // [nsys + 1] system s owns panels [sys_panel[s], sys_panel[s+1])
  unsigned short *jloc = nullptr;
  size_t cap_prow = 0, cap_pbeg = 0, cap_pcnt = 0, cap_pcol = 0, cap_pkey = 0, cap_jloc = 0, cap_sys_panel = 0;
  int npanel = 0, prow_n = -1;
  long long nbr_gen = 0, layout_gen = 0, prow_nbr_gen = -1, prow_layout_gen = -1;
  int panel_shape = -1;   // tuning family "panels": 0 (default) lists on, 1 off
  int *H_firstnbr = nullptr, *H_numnbrs = nullptr, *H_jlist = nullptr;
  double *H_val = nullptr;
  float *H_val32 = nullptr;
  long long Hcap = 0;
  double *Hdia_inv = nullptr, *b_s = nullptr, *b_t = nullptr;
  double *st = nullptr;            // [2][ncap]: s, t
  double *s_hist = nullptr, *t_hist = nullptr;  // [n][5]
  double *cgv = nullptr;           // CG vectors, see CGV_* below
  CgScalars *cgs = nullptr;        // [4 nsys]: s, t per system of the FP64 solve, then of the mixed inner solve
  double *partials = nullptr;      // per-workgroup partial dots, [8][cg_pstride]
  int cg_pstride = 0;
  unsigned int *counters = nullptr;

  // ---- batch of independent systems (one unless swr_set_batch)
  std::vector<int> sys_row;        // nsys + 1 local-row offsets (host)
  int nsys = 1, ntiles = 0;
  bool layout_dirty = true;
  int *row_sys = nullptr;          // [n] system of each local row
  int *tile_beg = nullptr;         // [ntiles + 1]
  int *tile_sys = nullptr;         // [ntiles]
  int *sys_tile = nullptr;         // [nsys + 1]
  // half-wave-per-row kernels over a batch: block b runs rows
  // [sblk_beg[b], sblk_beg[b+1]) (at most 8, one system), system s owns
  // blocks [sys_sblk[s], sys_sblk[s+1])
  int *sblk_beg = nullptr;         // [nsblk + 1]
  int *sys_sblk = nullptr;         // [nsys + 1]
  int nsblk = 0;
  // capacities of the layout arrays (grown, never shrunk)
  size_t cap_row_sys = 0, cap_tile_beg = 0, cap_tile_sys = 0, cap_sys_tile = 0, cap_sblk_beg = 0,
         cap_sys_sblk = 0, cap_sys_iters = 0, cap_cgs = 0, cap_qsum = 0;
  int *sys_iters = nullptr;        // [2 nsys] iterations of s, t per system (device)
  std::vector<int> h_sys_iters;    // host copy after each solve
  int *done_count = nullptr;       // systems finished in the running solve
  double *qsum = nullptr;          // [3 nsys] sum s, sum t, u per system
  int qeq_initialized = 0;
  int mixed = 0;                   // mixed-precision QEq
  void *mixw = nullptr;            // mixed-precision workspace
  size_t mixw_bytes = 0;
  int mixed_outer = 0;             // residual evaluations of the last mixed solve
  int mixed_last_inner = 0;        // iterations of the last inner FP32 solve (sizes the next launch batch)
  double mixed_rho0[2] = {0, 0};   // its first outer residual sqrt(u.r)/|b| (s, t)
  int imax = 200;
  // autotuned launch shapes (csrc/tune.cpp): SpMV shape of the FP64 CG [0],
  // of the FP32 inner solve [1] and of the FP64 residual [2], gatherer/
  // writer split of the far pass (-1: the default)
  int spmv_shape[3] = {-1, -1, -1};
  int far_shape = -1;
  std::string tune_report;         // JSON text of the last swr_autotune

  // ---- status
  DevStatus *status = nullptr;     // device
  DevStatus *status_host = nullptr;  // pinned
  unsigned int sticky_flags = 0;

  // ---- scratch for scans
  void *scan_tmp = nullptr;
  size_t scan_tmp_bytes = 0;
  int *tmp_i = nullptr;
  size_t tmp_i_cap = 0;

  // ---- per-kernel HIP-event timers (enabled by swr_ktimer_enable)
  bool ktime = false;
  struct KTimer {
    std::vector<hipEvent_t> ev;   // start/stop pairs
    int used = 0;                 // pairs recorded since the last read
    double ms = 0.0;
    long long count = 0;
  };
  KTimer kt[KT_COUNT];
};

extern const char *const swr_kt_names[KT_COUNT];
void swr_kt_begin(swr_system *S, int k);
void swr_kt_end(swr_system *S, int k);

// workspace array slots (each N doubles) for BO sections 1 and 3
enum {
  WS_DELTAP = 0, WS_DELTAP_BOC, WS_DELTA, WS_DELTA_E, WS_DELTA_VAL, WS_VLPEX, WS_NLP, WS_NLP_TEMP,
  WS_CLP, WS_DELTA_LP, WS_DDELTA_LP, WS_DELTA_LP_TEMP, WS_DDELTA_LP_TEMP, WS_COUNT
};
// CG vector slots (each ncap doubles, two right-hand sides interleaved per slot)
enum { CGV_R = 0, CGV_U, CGV_W, CGV_Z, CGV_SS, CGV_V, CGV_P, CGV_D0, CGV_D1, CGV_M, CGV_Q, CGV_COUNT };

// ---- helpers implemented in the .hip files
int swr_scan_i32_to_i64(swr_system *S, const int *in, long long *out, int n, long long *total_host);
int swr_scan_i32(swr_system *S, const int *in, int *out, int n, int *total_dev);
int swr_scan3_i32(swr_system *S, const int *const in[3], int *const out[3], int n, int *const total[3]);
int swr_ensure_tmp(swr_system *S, size_t bytes);

// Stage implementations (nbr.hip, lists.hip, bo.hip, qeq.hip, qeq_mixed.hip).
int swr_nbr_set_grid(swr_system *S, const double lo[3], const double hi[3], double cutghost, double cutneigh);
int swr_nbr_bbox(swr_system *S, double lo[3], double hi[3]);
int swr_nbr_build(swr_system *S, int local_only);
int swr_nbr_borders(swr_system *S, const double lo[3], const double hi[3], double cutghost);
int swr_nbr_forward_x(swr_system *S, const double lo[3], const double hi[3]);
int swr_nbr_nve(swr_system *S, double dt);
int swr_nbr_pbc(swr_system *S, const double lo[3], const double hi[3]);
int swr_nbr_ensure_atoms(swr_system *S, int Ncap);
int swr_lists_pack(swr_system *S);
int swr_lists_reset(swr_system *S);
int swr_lists_run(swr_system *S, int mode);
int swr_lists_estimate(swr_system *S);
int swr_lists_validate(swr_system *S);
int swr_lists_mark(swr_system *S, int max_sweeps);
int swr_mark_settle(swr_system *S);
int swr_after_sync(swr_system *S);
int swr_nonbonded_run(swr_system *S);
int swr_atom_energy_run(swr_system *S);
int swr_dbond_forces_run(swr_system *S);
int swr_hydrogen_bonds_run(swr_system *S);
int swr_valence_angles_run(swr_system *S);
int swr_torsion_angles_run(swr_system *S);
int swr_forces_run(swr_system *S);
int swr_nve_run(swr_system *S, int stage, double dtv, double dtf);
int swr_quiesce(swr_system *S);
constexpr int MARK_INNER = 8;         // bond_mark relaxation sweeps per launch
constexpr int MARK_MAX_LAUNCH = 128;  // launches of one relaxation at most
int swr_bo_run(swr_system *S);
int swr_qeq_build_H(swr_system *S);
int swr_qeq_spmv(swr_system *S, const double *xv, double *b);
int swr_qeq_cg_run(swr_system *S, int first_batch);
#define SWR_PIN_MIX 4096
// Grows the pinned iteration-count buffer to nsys systems (synchronises the
// stream before replacing it).
int swr_pin_iters_reserve(swr_system *S, int nsys);
int swr_qeq_calc_q(swr_system *S);
int swr_qeq_layout(swr_system *S);
int swr_qeq_mixed_solve(swr_system *S);
// autotuner timing hooks (qeq.hip, qeq_mixed.hip, lists.hip)
int swr_qeq_time_spmv(swr_system *S, int sh, int reps, float *ms);
int swr_qeq_time_mixed(swr_system *S, int which, int sh, int reps, float *ms);
// the no-allocation FP64 CG (qeq.hip): per-device scratch address, solve
int swr_noalloc_init(int device);
int swr_qeq_cg_noalloc_run(swr_system *S, const double *b, double *x, int *iters_s, int *iters_t);
int swr_status_to_rc(unsigned int flags);
int swr_lists_time_far(swr_system *S, int sh, int reps, float *ms);
// stable LDS radix sort of (key, value) pairs (nbr.hip)
int swr_radix_sort_i32(swr_system *S, const int *keys, const int *vals_in, int n, int nbits, int *vals_out);
// SpMV panels (panel.hip): row order + panel ranges when stale, then the
// column lists and local indices of the current H
int swr_panel_build(swr_system *S);

template <class T>
inline int swr_alloc(T **p, size_t count) {
  if (*p) { (void)hipFree(*p); *p = nullptr; }
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void **)p, count * sizeof(T));
  if (e != hipSuccess) {
    swr_set_error_msg(std::string("hipMalloc: ") + hipGetErrorString(e));
    *p = nullptr;
    return SWR_ERR_NOMEM;
  }
  return SWR_OK;
}

// Grows *p to at least count elements (a quarter more, so a slowly growing
// size does not reallocate every call); keeps it when it is large enough.
// hipFree waits for the whole device, so per-call buffers must not be
// freed and reallocated on every call: other systems' streams would stall.
template <class T>
inline int swr_reserve(T **p, size_t &cap, size_t count) {
  if (*p && count <= cap) return SWR_OK;
  const size_t want = count + count / 4 + 16;
  const int rc = swr_alloc(p, want);
  cap = rc ? 0 : want;
  return rc;
}

template <class T>
inline void swr_free(T *&p) {
  if (p) { (void)hipFree(p); p = nullptr; }
}

inline int div_up(long long a, long long b) { return (int)((a + b - 1) / b); }
