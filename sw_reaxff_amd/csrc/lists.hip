// lists.hip — atom pack (a2), Reset (a4), far list (a3), bonds + hbonds
// (a5, a6) and list validation (a7).
//
// The far/bond/hbond pass is one wave per atom i streaming i's Verlet
// candidates 64 at a time:
//   far   : d^2 <= nonb_cut^2, entry written with a wave-ballot compaction in
//           candidate order (write_reax_lists_c_para, pair_reaxc_sw64.c:276-331);
//   bond  : d <= bond_cut and uncorrected BO' >= bo_cut (BOp_single,
//           reaxc_forces_sunway.cpp:677-825; kernel reaxc_forces_sw64.c:1126-1280),
//           gathered in LDS, then sorted by nbr (my_bubble_sort, :31-75);
//   hbond : local H, far entry with d <= hbond_cut and p_hbond(j) == 2
//           (Init_Forces_noQEq_HB_Full_C, reaxc_forces_sw64.c:810-862).
// Far slots of atom i are [voff[i], voff[i+1]) — bounded by its Verlet
// count — so no separate counting pass is needed (the reference uses
// i*maxfar, pair_reaxc_sw64.c:256).
#include <type_traits>
#include <vector>

#include "swr_internal.h"
#include "swr_device.h"

namespace {

constexpr int LT = 256;                 // threads per block
constexpr int LW = LT / SWR_WAVE;       // waves per block
constexpr int MB = SWR_MAX_ATOM_BONDS;  // LDS bond slots per atom

// Orders this wave's LDS writes before its later LDS reads (each wave only
// touches its own LDS slice, so no workgroup barrier is needed).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ void k_pack(const int *__restrict__ tag, const int *__restrict__ type, const double *__restrict__ x,
                       const double *__restrict__ q, const DevParams *__restrict__ P, int N,
                       swr_atom_pack_t *__restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  swr_atom_pack_t a;
  a.orig_id = tag[i];
  a.type = P->map[type[i]];
  a.x[0] = x[3 * i]; a.x[1] = x[3 * i + 1]; a.x[2] = x[3 * i + 2];
  a.q = q[i];
  out[i] = a;
}

// Reset_Atoms flags / Reset_Neighbor_Lists inputs.
__global__ void k_reset_flags(const swr_atom_pack_t *__restrict__ P_atoms, const DevParams *__restrict__ P, int n,
                              int N, double hbond_cut, const int *__restrict__ num_hbonds, int *__restrict__ hflag,
                              int *__restrict__ hcap) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int t = P_atoms[i].type;
  int h = (i < n && hbond_cut > 0 && t >= 0 && P->p_hbond[t] == 1) ? 1 : 0;
  hflag[i] = h;
  hcap[i] = h ? num_hbonds[i] : 0;
}

__global__ void k_reset_apply(int n, int N, const int *__restrict__ hflag, const int *__restrict__ hpos,
                              const int *__restrict__ hstart, int *__restrict__ Hindex,
                              const int *__restrict__ bond_index, int *__restrict__ bond_end,
                              int *__restrict__ hbond_index, int *__restrict__ hbond_end, DevStatus *st,
                              int bond_cap, int hbond_cap) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    if (st->total_bonds > bond_cap) atomicOr(&st->flags, SWR_F_BOND_TOTAL);
    if (st->total_hbonds > hbond_cap) atomicOr(&st->flags, SWR_F_HBOND_TOTAL);
  }
  if (i >= N) return;
  bond_end[i] = bond_index[i];
  int H = hflag[i] ? hpos[i] : -1;
  Hindex[i] = H;
  if (H >= 0) hbond_index[H] = hbond_end[H] = hstart[i];
}

// Per atom, the far-list slots of its bond candidates (d <= bond_cut and
// inside the pair's BO' radius); the bond pass evaluates them.
constexpr int MCAND = 64;

// Streaming 16-byte store: the far list is written once per step and read
// back only sparsely, so it should not displace the gathered atoms in L2.
__device__ __forceinline__ void store_nt(double2 *dst, double2 v) {
  typedef double nd2 __attribute__((ext_vector_type(2)));
  nd2 t = {v.x, v.y};
  __builtin_nontemporal_store(t, reinterpret_cast<nd2 *>(dst));
}

// One far entry as four 16-byte pieces at dst (rel_box stays 0).
__device__ __forceinline__ void stage_far(double2 *dst, int j, int tj, int tagj, double d, double qj, double dx,
                                          double dy, double dz) {
  int4 *p = reinterpret_cast<int4 *>(dst);
  p[0] = make_int4(j, tj, tagj, 0);
  p[1] = make_int4(0, 0, __double2loint(d), __double2hiint(d));
  dst[2] = make_double2(qj, dx);
  dst[3] = make_double2(dy, dz);
}

// Uncorrected bond order of one pair (BOp_single, reaxc_forces_sunway.cpp:695-717).
struct BOpTerms { double BO, BO_s, pi, pi2, C12, C34, C56; };
__device__ __forceinline__ BOpTerms bop_terms(const DevParams *__restrict__ P, int ei, int ej, double d) {
  const DevTbp &tw = P->tbp[ei * SWR_MAX_ELEM + ej];
  BOpTerms t;
  if (P->r_s[ei] > 0.0 && P->r_s[ej] > 0.0) {
    t.C12 = tw.p_bo1 * pow(d / tw.r_s, tw.p_bo2);
    t.BO_s = (1.0 + P->bo_cut) * exp(t.C12);
  } else t.BO_s = t.C12 = 0.0;
  if (P->r_pi[ei] > 0.0 && P->r_pi[ej] > 0.0) {
    t.C34 = tw.p_bo3 * pow(d / tw.r_p, tw.p_bo4);
    t.pi = exp(t.C34);
  } else t.pi = t.C34 = 0.0;
  if (P->r_pi_pi[ei] > 0.0 && P->r_pi_pi[ej] > 0.0) {
    t.C56 = tw.p_bo5 * pow(d / tw.r_pp, tw.p_bo6);
    t.pi2 = exp(t.C56);
  } else t.pi2 = t.C56 = 0.0;
  t.BO = t.BO_s + t.pi + t.pi2;
  return t;
}

// Pass 1: one wave per atom streams its candidates (the Verlet list, or an
// existing far list when SRC_FAR) 64 at a time.
//   FAR  : write the far entries (write_reax_lists_c_para, pair_reaxc_sw64.c:276-331)
//   CAND : record bond candidates' far slots
//   HB   : 1 count, 2 write hbonds of local H (Init_Forces_noQEq_HB_Full_C,
//          reaxc_forces_sw64.c:810-862)
template <bool SRC_FAR, bool FAR, bool CAND, int HB>
__global__ void __launch_bounds__(LT)
k_far(int n, int N, const DevParams *__restrict__ P, const swr_atom_pack_t *__restrict__ atoms,
      const long long *__restrict__ voff, const int *__restrict__ vlist, swr_far_neighbor_t *__restrict__ far,
      int *__restrict__ far_index, int *__restrict__ far_end, int *__restrict__ cand, int *__restrict__ cand_count,
      const int *__restrict__ Hindex, const int *__restrict__ num_hbonds, const int *__restrict__ hbond_index,
      int *__restrict__ hbond_end, swr_hbond_t *__restrict__ hbonds, int *__restrict__ hb_count, int hbond_cap,
      DevStatus *st) {
  __shared__ double2 sfar[(FAR && !SRC_FAR) ? LW : 1][4 * SWR_WAVE];
  const int w = wave_in_block();
  const int blk = blockIdx.x;
  const int i = blk * LW + w;
  if (i >= N) return;
  const int lane = lane_id();
  const swr_atom_pack_t ai = atoms[i];
  const int ei = ai.type;
  const double nonb_cutsq = P->nonb_cutsq, bond_cut = P->bond_cut, hbond_cut = P->hbond_cut;
  long long cbeg, cend;
  if (SRC_FAR) { cbeg = far_index[i]; cend = far_end[i]; }
  else { cbeg = voff[i]; cend = voff[i + 1]; }
  const long long far0 = SRC_FAR ? cbeg : voff[i];
  int nfar = 0, ncand = 0;
  const bool doH = HB != 0 && (i < n) && ei >= 0 && P->p_hbond[ei] == 1 && hbond_cut > 0;
  const int H = (doH && HB == 2) ? Hindex[i] : -1;
  int htop = (doH && HB == 2) ? hbond_end[H] : 0;
  // capacity of this H's slots, never past the end of the list
  const int hlim = (doH && HB == 2) ? min(hbond_index[H] + num_hbonds[i], hbond_cap) : 0;
  const double *rbo = P->rbo + (ei >= 0 ? ei : 0) * SWR_MAX_ELEM;
  int *mycand = cand + (size_t)i * MCAND;

  int jnext = 0;
  if (!SRC_FAR && cbeg + lane < cend) jnext = vlist[cbeg + lane];
  for (long long base = cbeg; base < cend; base += SWR_WAVE) {
    const long long k = base + lane;
    const bool valid = k < cend;
    int j = 0, tj = -1, tagj = 0;
    double dx = 0, dy = 0, dz = 0, d = 0, qj = 0;
    bool infar = false;
    long long slot = 0;
    if (SRC_FAR) {
      if (valid) {
        const double2 *e = reinterpret_cast<const double2 *>(far + k);
        const int4 h0 = *reinterpret_cast<const int4 *>(far + k);
        const double2 e1 = e[1], e3 = e[3];
        const double2 e2 = e[2];
        j = h0.x; tj = h0.y; d = e1.y; dx = e2.y; dy = e3.x; dz = e3.y;
        infar = true;
        slot = k;
      }
    } else {
      j = jnext;
      if (k + SWR_WAVE < cend) jnext = vlist[k + SWR_WAVE];
      if (valid) {
        const swr_atom_pack_t aj = atoms[j];
        dx = aj.x[0] - ai.x[0]; dy = aj.x[1] - ai.x[1]; dz = aj.x[2] - ai.x[2];
        const double dsq = dx * dx + dy * dy + dz * dz;
        infar = dsq <= nonb_cutsq;
        if (infar) {
          d = sqrt(dsq);
          tj = aj.type; tagj = aj.orig_id; qj = aj.q;
        }
      }
      const unsigned long long m = __ballot(infar);
      const int r = ballot_rank(m);
      slot = far0 + nfar + r;
      if (FAR) {
        // Compact the wave's entries in LDS, then write them as contiguous
        // 16-byte pieces (lane t writes piece t, t+64, ...): full-line
        // stores instead of one 64-byte entry per lane.
        double2 *sw = sfar[w];
        if (infar) stage_far(sw + 4 * r, j, tj, tagj, d, qj, dx, dy, dz);
        wave_lds_sync();
        double2 *dst = reinterpret_cast<double2 *>(far + far0 + nfar);
        const int pieces = 4 * __popcll(m);
        for (int t = lane; t < pieces; t += SWR_WAVE) store_nt(dst + t, sw[t]);
        wave_lds_sync();
      }
      nfar += __popcll(m);
    }
    if (CAND) {
      const bool iscand = infar && ei >= 0 && tj >= 0 && d <= bond_cut && d <= rbo[tj];
      const unsigned long long cm = __ballot(iscand);
      if (iscand) {
        const int c = ncand + ballot_rank(cm);
        if (c < MCAND) mycand[c] = (int)slot;
      }
      ncand += __popcll(cm);
    }
    if (doH) {
      bool ish = false;
      if (infar && d <= hbond_cut) {
        const int tjj = SRC_FAR ? atoms[j].type : tj;
        ish = tjj >= 0 && P->p_hbond[tjj] == 2;
      }
      const unsigned long long hm = __ballot(ish);
      if (HB == 2 && ish) {
        const int pos = htop + ballot_rank(hm);
        if (pos < hlim) {
          swr_hbond_t hb;
          hb.nbr = j;
          hb.scl = 1;
          hb.ptr = far + slot;
          hbonds[pos] = hb;
        }
      }
      htop += __popcll(hm);
    }
  }
  if (lane == 0) {
    if (!SRC_FAR) {
      far_index[i] = (int)far0;
      far_end[i] = (int)(far0 + nfar);
    }
    if (CAND) {
      cand_count[i] = ncand;
      if (ncand > MCAND) atomicOr(&st->flags, SWR_F_BOND_LDS);
    }
    if (doH) {
      if (HB == 1) {
        hb_count[i] = htop;
      } else {
        if (htop > hlim) {
          atomicOr(&st->flags, SWR_F_HBOND_CAP);
          atomicMax(&st->need_hbond_cap, htop - hbond_index[H]);
        }
        hbond_end[H] = htop > hlim ? hlim : htop;
        hb_count[i] = htop - hbond_index[H];
      }
    }
  }
}

// Pass 1 for a Verlet-list source, split between two kinds of waves.
//
// A wave's vector-memory counter (vmcnt) counts loads and stores together
// and retires them in issue order, so a wave that writes its far entries
// and then gathers the next neighbours waits for its own stores to reach
// memory before every gather.  Here FG "gatherer" waves (one atom each)
// only load: per phase each takes one 64-candidate chunk of its atom's
// Verlet row, keeps d^2 <= nonb_cut^2 and stages the compacted 64-byte
// entries (plus hbond and bond-candidate records) in LDS, while FWR
// "writer" waves, which never load from global memory, stream the previous
// phase's staged entries out as contiguous 16-byte nontemporal stores.  One
// LDS barrier per phase; double-buffered staging.  Entry order and content
// are those of k_far (a far list is its Verlet row filtered in order).
// The split (FG gatherers, FWR writers per block) is a template parameter;
// the autotuner picks among the FarShape variants below.
template <int FG, int FWR>
struct FarStage {
  double2 far[2][FG][4 * SWR_WAVE];  // staged entries, 4 pieces each
  int2 hb[2][FG][SWR_WAVE];          // staged hbonds: (nbr, far slot)
  int cand[FG][MCAND];               // bond-candidate far slots of the atom
  double rbo[FG][SWR_MAX_ELEM];      // BO' radius row of the atom's element
  long long dst[2][FG];              // far slot of the phase's first entry
  int cnt[2][FG], hdst[2][FG], hcnt[2][FG];
  int nphase[FG];
};

// Gatherer/writer splits the autotuner chooses from.
template <int SH> struct FarShape;
template <> struct FarShape<SWR_FAR_G4W2> { static constexpr int FG = 4, FWR = 2; };
template <> struct FarShape<SWR_FAR_G4W4> { static constexpr int FG = 4, FWR = 4; };
template <> struct FarShape<SWR_FAR_G6W2> { static constexpr int FG = 6, FWR = 2; };
template <> struct FarShape<SWR_FAR_G8W4> { static constexpr int FG = 8, FWR = 4; };

template <class F>
void far_dispatch(int sh, F &&f) {
  switch (sh) {
    case SWR_FAR_G4W4: f(std::integral_constant<int, SWR_FAR_G4W4>()); break;
    case SWR_FAR_G6W2: f(std::integral_constant<int, SWR_FAR_G6W2>()); break;
    case SWR_FAR_G8W4: f(std::integral_constant<int, SWR_FAR_G8W4>()); break;
    default: f(std::integral_constant<int, SWR_FAR_G4W2>()); break;
  }
}

// A gatherer's neighbour record for one chunk: the id and the packed atom.
struct FarNbr {
  int j;
  swr_atom_pack_t a;
};

template <bool CAND, int HB, int FG, int FWR>
__global__ void __launch_bounds__((FG + FWR) * SWR_WAVE)
k_far_ws(int n, int N, const DevParams *__restrict__ P, const swr_atom_pack_t *__restrict__ atoms,
         const long long *__restrict__ voff, const int *__restrict__ vlist, swr_far_neighbor_t *__restrict__ far,
         int *__restrict__ far_index, int *__restrict__ far_end, int *__restrict__ cand,
         int *__restrict__ cand_count, const int *__restrict__ Hindex, const int *__restrict__ num_hbonds,
         const int *__restrict__ hbond_index, int *__restrict__ hbond_end, swr_hbond_t *__restrict__ hbonds,
         int *__restrict__ hb_count, int hbond_cap, DevStatus *st) {
  __shared__ FarStage<FG, FWR> S;
  const int w = __builtin_amdgcn_readfirstlane(wave_in_block());
  const int lane = lane_id();
  const bool gatherer = w < FG;

  // ---- gatherer set-up: the atom, its Verlet row, its element's tables
  const int g = gatherer ? w : 0;
  const int i = blockIdx.x * FG + g;
  const bool live = gatherer && i < N;
  swr_atom_pack_t ai = {};
  long long cbeg = 0, cend = 0;
  int ei = -1;
  if (live) {
    ai = atoms[i];
    cbeg = voff[i];
    cend = voff[i + 1];
    ei = __builtin_amdgcn_readfirstlane(ai.type);
  }
  const int nch = live ? (int)((cend - cbeg + SWR_WAVE - 1) / SWR_WAVE) : 0;
  const double nonb_cutsq = P->nonb_cutsq, bond_cut = P->bond_cut, hbond_cut = P->hbond_cut;
  unsigned hmask2 = 0;  // elements with p_hbond == 2 (acceptors)
  for (int e = 0; e < SWR_MAX_ELEM; e++) hmask2 |= (P->p_hbond[e] == 2 ? 1u : 0u) << e;
  const bool doH = HB != 0 && live && i < n && ei >= 0 && P->p_hbond[ei] == 1 && hbond_cut > 0;
  const int H = (doH && HB == 2) ? Hindex[i] : -1;
  int htop = (doH && HB == 2) ? hbond_end[H] : 0;
  const int hlim = (doH && HB == 2) ? min(hbond_index[H] + num_hbonds[i], hbond_cap) : 0;
  if (gatherer) {
    if (CAND && lane < SWR_MAX_ELEM) S.rbo[g][lane] = P->rbo[(ei >= 0 ? ei : 0) * SWR_MAX_ELEM + lane];
    if (lane == 0) {
      S.nphase[g] = nch;
      S.cnt[0][g] = S.cnt[1][g] = 0;
      S.hcnt[0][g] = S.hcnt[1][g] = 0;
    }
  }
  // Software pipeline without register copies (a copy of a loaded value
  // waits for the load): while chunk p is evaluated, chunk p+2's atoms and
  // chunk p+6's ids are requested, into three neighbour-record sets that
  // rotate with period 3 and six id registers that rotate with period 6.
  // An id load thus has four phases to arrive before its atoms are
  // requested (Verlet ids stream from HBM, whose read latency under the
  // far list's write stream is a few microseconds), an atom gather two.
  // Every phase of an active gatherer issues the same loads (indices clamped
  // to the row), and the gatherer, idle-gatherer and writer loops are
  // separate, so the compiler's wait counts let the loads stay in flight.
  const long long clast = cend - 1;
  FarNbr X = {}, Y = {}, Z = {};
  int I0 = 0, I1 = 0, I2 = 0, I3 = 0, I4 = 0, I5 = 0;
  if (nch > 0) {
    I0 = vlist[min(cbeg + lane, clast)];
    I1 = vlist[min(cbeg + SWR_WAVE + lane, clast)];
    I2 = vlist[min(cbeg + 2 * SWR_WAVE + lane, clast)];
    I3 = vlist[min(cbeg + 3 * SWR_WAVE + lane, clast)];
    I4 = vlist[min(cbeg + 4 * SWR_WAVE + lane, clast)];
    I5 = vlist[min(cbeg + 5 * SWR_WAVE + lane, clast)];
    X.a = atoms[I0];
    X.j = I0;
    Y.a = atoms[I1];
    Y.j = I1;
  }
  lds_barrier();
  int nph = 0;
  for (int k = 0; k < FG; k++) nph = max(nph, S.nphase[k]);
  // phases 0..nph (the writers drain phase p-1 in phase p); identical in
  // every wave of the block
  const int nphr = __builtin_amdgcn_readfirstlane(nph + 1);

  const long long far0 = cbeg;
  int nfar = 0, ncand = 0;

  // gatherer phase: evaluate chunk p (cur), request chunk p+2's atoms (ids
  // jaddr) into far2 and then chunk p+6's ids into jfree (in this order:
  // waiting for a chunk's atoms then never waits for the ids requested
  // after them)
  auto gather = [&](int p, const FarNbr &cur, FarNbr &far2, const int &jaddr, int &jfree) {
    const int buf = p & 1;
    const long long k = cbeg + (long long)p * SWR_WAVE + lane;
    far2.a = atoms[jaddr];
    far2.j = jaddr;
    jfree = vlist[min(k + 6 * SWR_WAVE, clast)];
    if (p < nch) {
      const bool valid = k < cend;
      double dx = 0, dy = 0, dz = 0, d = 0;
      bool infar = false;
      if (valid) {
        dx = cur.a.x[0] - ai.x[0]; dy = cur.a.x[1] - ai.x[1]; dz = cur.a.x[2] - ai.x[2];
        const double dsq = dx * dx + dy * dy + dz * dz;
        infar = dsq <= nonb_cutsq;
        if (infar) d = sqrt(dsq);
      }
      const int tj = infar ? cur.a.type : -1;
      const unsigned long long m = __ballot(infar);
      const int r = ballot_rank(m);
      const long long slot = far0 + nfar + r;
      if (infar) stage_far(S.far[buf][g] + 4 * r, cur.j, tj, cur.a.orig_id, d, cur.a.q, dx, dy, dz);
      if (lane == 0) {
        S.dst[buf][g] = far0 + nfar;
        S.cnt[buf][g] = __popcll(m);
      }
      if (CAND) {
        const bool iscand = infar && ei >= 0 && tj >= 0 && d <= bond_cut && d <= S.rbo[g][tj];
        const unsigned long long cm = __ballot(iscand);
        if (iscand) {
          const int c = ncand + ballot_rank(cm);
          if (c < MCAND) S.cand[g][c] = (int)slot;
        }
        ncand += __popcll(cm);
      }
      if (doH) {
        const bool ish = infar && d <= hbond_cut && tj >= 0 && ((hmask2 >> tj) & 1u);
        const unsigned long long hm = __ballot(ish);
        if (HB == 2) {
          const int hr = ballot_rank(hm);
          if (ish && htop + hr < hlim) S.hb[buf][g][hr] = make_int2(cur.j, (int)slot);
          if (lane == 0) {
            S.hdst[buf][g] = htop;
            S.hcnt[buf][g] = max(0, min(__popcll(hm), hlim - htop));
          }
        }
        htop += __popcll(hm);
      } else if (HB == 2 && lane == 0) {
        S.hcnt[buf][g] = 0;
      }
      nfar += __popcll(m);
    } else if (lane == 0) {
      S.cnt[buf][g] = 0;
      S.hcnt[buf][g] = 0;
    }
    lds_barrier();
  };

  if (gatherer && nch > 0) {
    // six phases per trip (the period of the id ring): a loop with one exit,
    // then the last nphr % 6 phases in straight-line code continuing the
    // rotation (early exits from the loop make the compiler's wait counts
    // at its header conservative); nphr is wave-uniform, so every wave of
    // the block passes nphr barriers
    int p = 0;
    for (; p + 6 <= nphr; p += 6) {
      gather(p, X, Z, I2, I0);
      gather(p + 1, Y, X, I3, I1);
      gather(p + 2, Z, Y, I4, I2);
      gather(p + 3, X, Z, I5, I3);
      gather(p + 4, Y, X, I0, I4);
      gather(p + 5, Z, Y, I1, I5);
    }
    if (p < nphr) gather(p, X, Z, I2, I0);
    if (p + 1 < nphr) gather(p + 1, Y, X, I3, I1);
    if (p + 2 < nphr) gather(p + 2, Z, Y, I4, I2);
    if (p + 3 < nphr) gather(p + 3, X, Z, I5, I3);
    if (p + 4 < nphr) gather(p + 4, Y, X, I0, I4);
  } else if (gatherer) {
    for (int p = 0; p < nphr; p++) lds_barrier();
  } else {
    // writer: in phase p, drain phase p-1's staging of its gatherers
    for (int p = 0; p < nphr; p++) {
      if (p > 0) {
        const int pb = (p - 1) & 1;
        for (int gg = w - FG; gg < FG; gg += FWR) {
          const int cnt = S.cnt[pb][gg];
          if (cnt > 0) {
            double2 *dstp = reinterpret_cast<double2 *>(far + S.dst[pb][gg]);
            const double2 *src = S.far[pb][gg];
            for (int t = lane; t < 4 * cnt; t += SWR_WAVE) store_nt(dstp + t, src[t]);
          }
          if (HB == 2) {
            const int hc = S.hcnt[pb][gg];
            if (lane < hc) {
              const int2 e = S.hb[pb][gg][lane];
              swr_hbond_t hb;
              hb.nbr = e.x;
              hb.scl = 1;
              hb.ptr = far + e.y;
              hbonds[S.hdst[pb][gg] + lane] = hb;
            }
          }
        }
      }
      lds_barrier();
    }
  }
  if (!live) return;
  // per-atom results (the gatherer's only global stores)
  if (CAND) {
    wave_lds_sync();
    const int nc = min(ncand, MCAND);
    for (int t = lane; t < nc; t += SWR_WAVE) cand[(size_t)i * MCAND + t] = S.cand[g][t];
  }
  if (lane == 0) {
    far_index[i] = (int)far0;
    far_end[i] = (int)(far0 + nfar);
    if (CAND) {
      cand_count[i] = ncand;
      if (ncand > MCAND) atomicOr(&st->flags, SWR_F_BOND_LDS);
    }
    if (doH) {
      if (HB == 1) {
        hb_count[i] = htop;
      } else {
        if (htop > hlim) {
          atomicOr(&st->flags, SWR_F_HBOND_CAP);
          atomicMax(&st->need_hbond_cap, htop - hbond_index[H]);
        }
        hbond_end[H] = htop > hlim ? hlim : htop;
        hb_count[i] = htop - hbond_index[H];
      }
    }
  }
}

// Pass 2: BL lanes per atom, SWR_WAVE / BL atoms per wave, evaluate the
// atom's bond candidates (one per lane, BL at a time), keep BO' >= bo_cut in
// candidate (= far-list) order, sort the bonds by nbr (my_bubble_sort,
// reaxc_forces_sw64.c:31-75) and write the bond records, the per-atom sums
// and BO section 1.  An atom has a handful of candidates (about 3 in water,
// 6 in water/methanol clusters), so wider groups left most lanes idle; a
// bond's LDS slot keeps only its far slot and neighbour, and the write pass
// reads the far entry again.  The per-atom sums run in slot order (one
// lane of the group adds the group's values in order through shuffles).
constexpr int BL = 8;   // lanes per atom in k_bonds

template <bool COUNT_ONLY>
__global__ void __launch_bounds__(LT)
k_bonds(int n, int N, const DevParams *__restrict__ P, const swr_atom_pack_t *__restrict__ atoms,
        const swr_far_neighbor_t *__restrict__ far, const int *__restrict__ cand,
        const int *__restrict__ cand_count, const int *__restrict__ num_bonds, const int *__restrict__ bond_index,
        int *__restrict__ bond_end, swr_bond_t *__restrict__ bonds, int *__restrict__ bond_nbr,
        swr_bond_order_t *__restrict__ bo_data,
        double *__restrict__ BO, double *__restrict__ BOpi, double *__restrict__ Cdbo, double *__restrict__ Cdbopi,
        double *__restrict__ Cdbopi2, double *__restrict__ bo_dboc, double *__restrict__ dDeltap_self,
        int *__restrict__ bond_mark, double *__restrict__ Deltap, double *__restrict__ Deltap_boc,
        int *__restrict__ bond_count, int bond_cap, DevStatus *st) {
  constexpr int APW = SWR_WAVE / BL;   // atoms per wave
  __shared__ int sfar[APW * LW][MB];   // far slot of the atom's m-th bond (candidate order)
  __shared__ int snbr[APW * LW][MB];
  const int w = wave_in_block();
  const int lane = lane_id();
  const int grp = lane / BL, sl = lane % BL;
  const int slot_a = APW * w + grp;  // this atom's LDS row
  const int i0 = (xcd_block(gridDim.x) * LW + w) * APW;
  if (i0 >= N) return;               // every atom of the wave past the end
  const int i = i0 + grp;
  const bool live = i < N;
  const int ei = live ? atoms[i].type : -1;
  const double bo_cut = P->bo_cut;
  const int ncand = live ? min(cand_count[i], MCAND) : 0;
  int ncmax = ncand;                 // uniform loop bound over the wave's atoms
#pragma unroll
  for (int o = BL; o < SWR_WAVE; o <<= 1) ncmax = max(ncmax, __shfl_xor(ncmax, o, SWR_WAVE));
  const int *mycand = cand + (size_t)(live ? i : 0) * MCAND;
  const unsigned long long gmask = (1ull << BL) - 1ull;
  const unsigned long long below = (1ull << sl) - 1ull;
  int nb = 0;
  for (int c0 = 0; c0 < ncmax; c0 += BL) {
    const int c = c0 + sl;
    bool isb = false;
    int fs = 0, jn = 0;
    if (c < ncand && ei >= 0) {
      fs = mycand[c];
      const swr_far_neighbor_t e = far[fs];
      const BOpTerms t = bop_terms(P, ei, e.type, e.d);
      if (t.BO >= bo_cut) {
        isb = true;
        jn = e.nbr;
      }
    }
    const unsigned long long bm = (__ballot(isb) >> (BL * grp)) & gmask;
    if (!COUNT_ONLY && isb) {
      const int slot = nb + __popcll(bm & below);
      if (slot < MB) {
        sfar[slot_a][slot] = fs;
        snbr[slot_a][slot] = jn;
      }
    }
    nb += __popcll(bm);
  }
  if (live && sl == 0) bond_count[i] = nb;
  if (COUNT_ONLY) return;
  const int cap = live ? num_bonds[i] : 0;
  const int start = live ? bond_index[i] : 0;
  const bool over = live && (nb > MB || nb > cap || start + nb > bond_cap);
  if (over && sl == 0) {
    atomicOr(&st->flags, nb > MB ? (SWR_F_BOND_LDS | SWR_F_BOND_CAP) : SWR_F_BOND_CAP);
    atomicMax(&st->need_bond_cap, nb);
    bond_end[i] = start;
  }
  const bool write = live && !over;
  const int nbw = write ? nb : 0;
  int nbmax = nbw;
#pragma unroll
  for (int o = BL; o < SWR_WAVE; o <<= 1) nbmax = max(nbmax, __shfl_xor(nbmax, o, SWR_WAVE));
  wave_lds_sync();
  // bo_dboc[i][0] and dDeltap_self[i] accumulate in candidate order
  // (reaxc_forces_sw64.c:1253,1262), starting from Reset's zeros.
  double tot = 0.0, s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int m0 = 0; m0 < nbmax; m0 += BL) {
    const int m = m0 + sl;
    double bm_ = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0;
    if (m < nbw) {
      const swr_far_neighbor_t e = far[sfar[slot_a][m]];
      const int jn = snbr[slot_a][m];
      const BOpTerms t = bop_terms(P, ei, e.type, e.d);
      const DevTbp &tw = P->tbp[ei * SWR_MAX_ELEM + e.type];
      const double r2 = e.d * e.d;
      const double Cs = tw.p_bo2 * t.C12 / r2, Cp = tw.p_bo4 * t.C34 / r2, Cpp = tw.p_bo6 * t.C56 / r2;
      const double dv[3] = {e.dvec[0], e.dvec[1], e.dvec[2]};
      const double cdb = -(t.BO_s * Cs + t.pi * Cp + t.pi2 * Cpp);
      swr_bond_order_t o;
      o.BO_s = t.BO_s - bo_cut;
      for (int c = 0; c < 3; c++) {
        o.dln_BOp_s[c] = -t.BO_s * Cs * dv[c];
        o.dln_BOp_pi[c] = -t.pi * Cp * dv[c];
        o.dln_BOp_pi2[c] = -t.pi2 * Cpp * dv[c];
        o.dBOp[c] = cdb * dv[c];
      }
      o.C1dbo = o.C2dbo = o.C3dbo = 0.0;
      o.C1dbopi = o.C2dbopi = o.C3dbopi = o.C4dbopi = 0.0;
      o.C1dbopi2 = o.C2dbopi2 = o.C3dbopi2 = o.C4dbopi2 = 0.0;
      o.CdboReduction = nullptr;
      int rank = 0;
      for (int q = 0; q < nbw; q++) rank += (snbr[slot_a][q] < jn) ? 1 : 0;
      const int p = start + rank;
      swr_bond_t b;
      b.d = e.d; b.dvec[0] = dv[0]; b.dvec[1] = dv[1]; b.dvec[2] = dv[2];
      b.nbr = jn;
      b.dbond_index = start + m;
      b.rel_box[0] = b.rel_box[1] = b.rel_box[2] = 0;
      b.sym_index = 0;
      bonds[p] = b;
      bond_nbr[p] = jn;
      bo_data[p] = o;
      BO[p] = t.BO - bo_cut;
      BOpi[2 * (size_t)p] = t.pi;
      BOpi[2 * (size_t)p + 1] = t.pi2;
      Cdbo[p] = 0.0; Cdbopi[p] = 0.0; Cdbopi2[p] = 0.0;
      bm_ = t.BO - bo_cut;
      g0 = o.dBOp[0]; g1 = o.dBOp[1]; g2 = o.dBOp[2];
    }
    // the group's slots m0 .. m0 + BL - 1 in order, onto its first lane
    for (int k = 0; k < BL; k++) {
      const int src = grp * BL + k;
      const double vb = __shfl(bm_, src, SWR_WAVE), v0 = __shfl(g0, src, SWR_WAVE);
      const double v1 = __shfl(g1, src, SWR_WAVE), v2 = __shfl(g2, src, SWR_WAVE);
      if (m0 + k < nbw) {
        tot += vb; s0 += v0; s1 += v1; s2 += v2;
      }
    }
  }
  if (write && sl == 0) {
    bond_end[i] = start + nb;
    bo_dboc[2 * (size_t)i] = tot;
    bo_dboc[2 * (size_t)i + 1] = 0.0;
    dDeltap_self[3 * (size_t)i] = s0;
    dDeltap_self[3 * (size_t)i + 1] = s1;
    dDeltap_self[3 * (size_t)i + 2] = s2;
    bond_mark[i] = (i < n) ? 0 : 1000;
    // BO section 1 (reaxc_bond_orders_sw64.c:56-64), computed here since it
    // only needs this atom's bond-order sum.
    if (ei >= 0) {
      Deltap[i] = tot - P->valency[ei];
      Deltap_boc[i] = tot - P->valency_val[ei];
    }
  }
}

// Counts come from the list pass (bond_count/hb_count hold the true
// per-atom totals even when a capacity overflowed, so the next capacities
// are large enough to retry the step).
__global__ void k_validate(int n, int N, const int *__restrict__ bond_index, const int *__restrict__ bond_count,
                           int *__restrict__ num_bonds, const int *__restrict__ Hindex,
                           const int *__restrict__ hb_count, int *__restrict__ num_hbonds, double saferzone,
                           int bond_total, DevStatus *st) {
  __shared__ int sb[8], sh[8];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int tb = 0, th = 0;
  if (i < N) {
    const int c = bond_count[i];
    tb = max(c * 2, 25);
    num_bonds[i] = tb;
    const int comp = (i < N - 1) ? bond_index[i + 1] : bond_total;
    if (bond_index[i] + c > comp) atomicOr(&st->flags, SWR_F_BOND_CAP);
    if (i < n) {
      const int H = Hindex[i];
      if (H > -1) {
        const double v = hb_count[i] * saferzone;
        th = (int)(v > 25.0 ? v : 25.0);
        num_hbonds[i] = th;
      }
    }
  }
  // totals of the new capacities: the next step's Reset sizes its lists from
  // them without a host round trip (integer sums: order-independent)
  tb = wave_sum(tb);
  th = wave_sum(th);
  if (lane_id() == 0) { sb[wave_in_block()] = tb; sh[wave_in_block()] = th; }
  __syncthreads();
  if (threadIdx.x == 0) {
    int a = 0, b = 0;
    for (int w = 0; w < (int)blockDim.x / SWR_WAVE; w++) { a += sb[w]; b += sh[w]; }
    atomicAdd(&st->next_total_bonds, a);
    atomicAdd(&st->next_total_hbonds, b);
  }
}

// sort_bonds_C (reaxc_forces_sw64.c:932-971): one in-order pass
// bond_mark[i] = min(bond_mark[i], bond_mark[j]+1).  Its result for ghost i
// is min(1000, 1 + min over bonded j < i of the result at j) (locals stay 0,
// later ghosts still hold 1000); Jacobi sweeps over that recursion reach the
// same fixed point.  flags[0..2]: sweep k reads flags[(k+2)%3] (changed in
// sweep k-1), sets flags[k%3], clears flags[(k+1)%3].  flags[3] turns 1 once
// a sweep finds that the previous one changed nothing; later sweeps skip.
// In-place relaxation of the marks of ghost atoms (local atoms are 0 and
// never change).  A ghost g takes min(mark[g], mark[j] + 1) over its bonds to
// lower-index atoms j, `inner` times per launch, reading neighbours' current
// values (relaxed atomic loads, so other workgroups' updates are seen as they
// land).  Every value written is >= the value of the reference's sequential
// pass and the updates only follow j < g, so the marks settle on exactly that
// pass's result whatever the update order; a launch in which no mark changed
// has reached it.  chg[launch] is set when the launch changed a mark; a launch
// after an unchanged one does nothing.
__global__ void k_bond_mark(int n, int N, int launch, int inner, const int *__restrict__ bond_index,
                            const int *__restrict__ bond_end, const int *__restrict__ bond_nbr,
                            int *__restrict__ mark, int *__restrict__ chg) {
  if (launch > 0 && __hip_atomic_load(&chg[launch - 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) return;
  const int g = n + blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= N) return;
  const int p0 = bond_index[g], p1 = bond_end[g];
  int cur = __hip_atomic_load(&mark[g], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  bool changed = false;
  for (int it = 0; it < inner; it++) {
    int best = cur;
    for (int p = p0; p < p1; p++) {
      const int j = bond_nbr[p];
      if (j < g) {
        const int v = __hip_atomic_load(&mark[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1;
        if (v < best) best = v;
      }
    }
    if (best < cur) {
      cur = best;
      changed = true;
      __hip_atomic_store(&mark[g], best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  if (changed) atomicOr(&chg[launch], 1);
}
__global__ void k_estimate(int n, int N, const int *__restrict__ bond_count, const int *__restrict__ hb_count,
                           const swr_atom_pack_t *__restrict__ atoms, const DevParams *__restrict__ P,
                           double saferzone, int *__restrict__ num_bonds, int *__restrict__ num_hbonds) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  if (i < n) {
    num_bonds[i] = max(bond_count[i] * 2, 25);
    const int t = atoms[i].type;
    const bool isH = t >= 0 && P->p_hbond[t] == 1 && P->hbond_cut > 0;
    const double v = (isH ? hb_count[i] : 0) * saferzone;
    num_hbonds[i] = (int)(v > 25.0 ? v : 25.0);
  } else {
    num_bonds[i] = 25;
    num_hbonds[i] = 0;
  }
}

__global__ void k_mark_init(int *__restrict__ chg, int nchg) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nchg) chg[i] = 0;
}
__global__ void k_set_flag(DevStatus *st, unsigned int f) { atomicOr(&st->flags, f); }

// Reports whether the sweeps converged and at which sweep (host reads it).
}  // namespace

// ---------------------------------------------------------------- host side
int swr_lists_pack(swr_system *S) {
  hipLaunchKernelGGL(k_pack, dim3(div_up(S->N, 256)), dim3(256), 0, S->stream, S->tag, S->type, S->x, S->q, S->dp,
                     S->N, S->packed);
  HIP_CHECK(hipGetLastError());
  return SWR_OK;
}

int swr_lists_alloc_bonds(swr_system *S, int bond_cap, int hbond_cap);

int swr_lists_reset(swr_system *S) {
  const int n = S->n, N = S->N;
  int rc = swr_ensure_tmp(S, sizeof(int) * (4 * (size_t)N + 8));
  if (rc) return rc;
  int *hflag = S->tmp_i, *hcap = hflag + N + 1, *hpos = hcap + N + 1, *hstart = hpos + N + 1;
  swr_kt_begin(S, KT_RESET);
  // Reset_Workspace (reaxc_reset_tools_sunway.cpp:122-131) zeroes fCdDelta,
  // which the non-bonded pass accumulates into
  HIP_CHECK(hipMemsetAsync(S->fcd, 0, sizeof(double) * 4 * (size_t)N, S->stream));
  hipLaunchKernelGGL(k_reset_flags, dim3(div_up(N, 256)), dim3(256), 0, S->stream, S->packed, S->dp, n, N,
                     S->ctl.hbond_cut, S->num_hbonds, hflag, hcap);
  {  // the three prefix sums of Reset in one set of launches
    const int *in[3] = {hflag, hcap, S->num_bonds};
    int *out[3] = {hpos, hstart, S->bond_index};
    int *tot[3] = {&S->status->numH, &S->status->total_hbonds, &S->status->total_bonds};
    if ((rc = swr_scan3_i32(S, in, out, N, tot))) return rc;
  }
  hipLaunchKernelGGL(k_reset_apply, dim3(div_up(N, 256)), dim3(256), 0, S->stream, n, N, hflag, hpos, hstart,
                     S->Hindex, S->bond_index, S->bond_end, S->hbond_index, S->hbond_end, S->status, S->bond_cap,
                     S->hbond_cap);
  swr_kt_end(S, KT_RESET);
  HIP_CHECK(hipGetLastError());
  return SWR_OK;
}

int swr_lists_run(swr_system *S, int mode) {
  {  // k_bonds rewrites bond_mark: finish the last relaxation first
    int rc = swr_quiesce(S);
    if (rc) return rc;
  }
  const int N = S->N;
  const int nblk = div_up(N, LW);
  double *ws = S->ws;
  int rc = SWR_OK;
  if (mode != 0 && (size_t)N * MCAND > S->cand_cap) {
    if ((rc = swr_alloc(&S->cand, (size_t)S->Ncap * MCAND)) || (rc = swr_alloc(&S->cand_count, S->Ncap))) return rc;
    S->cand_cap = (size_t)S->Ncap * MCAND;
  }
  int *hb_count = S->counts + S->Ncap;
#define FAR_ARGS(k)                                                                                              \
  hipLaunchKernelGGL(k, dim3(nblk), dim3(LT), 0, S->stream, S->n, N, S->dp, S->packed, S->voff, S->vlist, S->far, \
                     S->far_index, S->far_end, S->cand, S->cand_count, S->Hindex, S->num_hbonds, S->hbond_index,  \
                     S->hbond_end, S->hbonds, hb_count, S->hbond_cap, S->status)
#define BOND_ARGS(k)                                                                                              \
  hipLaunchKernelGGL(k, dim3(div_up(N, (SWR_WAVE / BL) * LW)), dim3(LT), 0, S->stream, S->n, N, S->dp, S->packed,     \
                     S->far, S->cand,                                                                              \
                     S->cand_count, S->num_bonds, S->bond_index, S->bond_end, S->bonds, S->bond_nbr, S->bo_data,    \
                     S->BO, S->BOpi,                                                                               \
                     S->Cdbo, S->Cdbopi, S->Cdbopi2, S->bo_dboc, S->dDeltap_self, S->bond_mark,                    \
                     ws + (size_t)WS_DELTAP * S->Ncap, ws + (size_t)WS_DELTAP_BOC * S->Ncap, S->counts,            \
                     S->bond_cap, S->status)
  swr_kt_begin(S, KT_FAR);
  if (mode == 2) {
    FAR_ARGS((k_far<true, false, true, 2>));
  } else {
    far_dispatch(S->far_shape, [&](auto shc) {
      constexpr int FG = FarShape<decltype(shc)::value>::FG, FWR = FarShape<decltype(shc)::value>::FWR;
      const dim3 gb(div_up(N, FG)), tb((FG + FWR) * SWR_WAVE);
#define FARWS_ARGS(k)                                                                                            \
  hipLaunchKernelGGL(k, gb, tb, 0, S->stream, S->n, N, S->dp, S->packed, S->voff, S->vlist, S->far, S->far_index,    \
                     S->far_end, S->cand, S->cand_count, S->Hindex, S->num_hbonds, S->hbond_index, S->hbond_end,     \
                     S->hbonds, hb_count, S->hbond_cap, S->status)
      if (mode == 0) FARWS_ARGS((k_far_ws<false, 0, FG, FWR>));
      else if (mode == 1) FARWS_ARGS((k_far_ws<true, 2, FG, FWR>));
      else FARWS_ARGS((k_far_ws<true, 1, FG, FWR>));
    });
  }
  swr_kt_end(S, KT_FAR);
  if (mode != 0) {
    swr_kt_begin(S, KT_BONDS);
    if (mode == 3) BOND_ARGS((k_bonds<true>));
    else BOND_ARGS((k_bonds<false>));
    swr_kt_end(S, KT_BONDS);
  }
#undef FAR_ARGS
#undef FARWS_ARGS
#undef BOND_ARGS
  HIP_CHECK(hipGetLastError());
  return SWR_OK;
}

// Autotuner hook: the average time of `reps` launches of the far pass (the
// fused far + bond-candidate + hbond form) in gatherer/writer shape sh.
// Each launch follows a Reset, as in a step (the pass appends to the hbond
// lists Reset starts); the caller re-runs the pair stages afterwards.
int swr_lists_time_far(swr_system *S, int sh, int reps, float *ms) {
  int rc = swr_quiesce(S);
  if (rc) return rc;
  const int N = S->N;
  if (N <= 0 || !S->far || !S->packed) return SWR_ERR_STATE;
  if ((size_t)N * MCAND > S->cand_cap) {
    if ((rc = swr_alloc(&S->cand, (size_t)S->Ncap * MCAND)) || (rc = swr_alloc(&S->cand_count, S->Ncap))) return rc;
    S->cand_cap = (size_t)S->Ncap * MCAND;
  }
  int *hb_count = S->counts + S->Ncap;
  std::vector<hipEvent_t> ev(2 * (size_t)std::max(reps, 1));
  for (auto &e : ev) HIP_CHECK(hipEventCreate(&e));
  for (int r = 0; r < reps; r++) {
    if ((rc = swr_lists_reset(S))) return rc;
    (void)hipEventRecord(ev[2 * r], S->stream);
    far_dispatch(sh, [&](auto shc) {
      constexpr int FG = FarShape<decltype(shc)::value>::FG, FWR = FarShape<decltype(shc)::value>::FWR;
      hipLaunchKernelGGL((k_far_ws<true, 2, FG, FWR>), dim3(div_up(N, FG)), dim3((FG + FWR) * SWR_WAVE), 0,
                         S->stream, S->n, N, S->dp, S->packed, S->voff, S->vlist, S->far, S->far_index, S->far_end,
                         S->cand, S->cand_count, S->Hindex, S->num_hbonds, S->hbond_index, S->hbond_end, S->hbonds,
                         hb_count, S->hbond_cap, S->status);
    });
    (void)hipEventRecord(ev[2 * r + 1], S->stream);
  }
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(S->stream));
  double tot = 0.0;
  for (int r = 0; r < reps; r++) {
    float t = 0.f;
    HIP_CHECK(hipEventElapsedTime(&t, ev[2 * r], ev[2 * r + 1]));
    tot += t;
  }
  for (auto &e : ev) (void)hipEventDestroy(e);
  *ms = (float)(tot / std::max(reps, 1));
  return SWR_OK;
}

int swr_lists_estimate(swr_system *S) {
  int rc = swr_lists_run(S, 3);
  if (rc) return rc;
  hipLaunchKernelGGL(k_estimate, dim3(div_up(S->N, 256)), dim3(256), 0, S->stream, S->n, S->N, S->counts,
                     S->counts + S->Ncap, S->packed, S->dp, S->ctl.saferzone, S->num_bonds, S->num_hbonds);
  HIP_CHECK(hipGetLastError());
  return SWR_OK;
}

// Validate_Lists: next capacities and overflow flags; the status block
// (with the capacity totals Reset needs next step) is copied back
// asynchronously and read at the host's next sync.
int swr_lists_validate(swr_system *S) {
  const int n = S->n, N = S->N;
  swr_kt_begin(S, KT_VALIDATE);
  HIP_CHECK(hipMemsetAsync(&S->status->next_total_bonds, 0, 2 * sizeof(int), S->stream));
  hipLaunchKernelGGL(k_validate, dim3(div_up(N, 256)), dim3(256), 0, S->stream, n, N, S->bond_index, S->counts,
                     S->num_bonds, S->Hindex, S->counts + S->Ncap, S->num_hbonds, S->ctl.saferzone, S->bond_cap,
                     S->status);
  // the capacity totals travel back with the next sync (next step's Reset)
  HIP_CHECK(hipMemcpyAsync(S->status_host, S->status, sizeof(DevStatus), hipMemcpyDeviceToHost, S->stream));
  S->status_copy_queued = true;
  S->next_totals_known = false;
  swr_kt_end(S, KT_VALIDATE);
  HIP_CHECK(hipGetLastError());
  return SWR_OK;
}

// bond_mark of sort_bonds_C (reaxc_forces_sw64.c:932-971): one pass over
// the atoms in index order, each taking min(mark[i], mark[j] + 1) over its
// bonds with the values current at that point.  Local atoms start at 0 and
// ghosts at 1000, so only ghosts change, and ghost i ends at
//   min(1000, min over bonded j < i of final[j] + 1, min over j > i of init[j] + 1).
// The sweeps below evaluate that recursion over the j < i graph (Jacobi
// sweeps: sweep s fixes every ghost whose dependency chain has depth <= s)
// until no mark changes; the host reads the convergence flag between batches.
// Enqueues the relaxation (last step's launch count + 2, no host wait) and a
// copy of the per-launch change flags; swr_mark_settle checks them at the next
// sync and runs further launches only if the last one still changed a mark.
int swr_lists_mark(swr_system *S, int max_sweeps) {
  const int n = S->n, N = S->N;
  S->mark_launched = 0;
  if (N <= n) return SWR_OK;
  if (!S->mark_chg) {
    int rc = swr_alloc(&S->mark_chg, MARK_MAX_LAUNCH);
    if (rc) return rc;
    if (hipHostMalloc((void **)&S->mark_chg_host, sizeof(int) * MARK_MAX_LAUNCH, 0) != hipSuccess) {
      swr_set_error_msg("hipHostMalloc failed");
      return SWR_ERR_NOMEM;
    }
  }
  const int max_launch = std::min(MARK_MAX_LAUNCH, div_up(max_sweeps, MARK_INNER));
  const int L = std::min(max_launch, std::max(2, div_up(S->mark_depth_hint, MARK_INNER) + 2));
  swr_kt_begin(S, KT_MARK);
  hipLaunchKernelGGL(k_mark_init, dim3(1), dim3(MARK_MAX_LAUNCH), 0, S->stream, S->mark_chg, MARK_MAX_LAUNCH);
  for (int k = 0; k < L; k++)
    hipLaunchKernelGGL(k_bond_mark, dim3(div_up(N - n, 256)), dim3(256), 0, S->stream, n, N, k, MARK_INNER,
                       S->bond_index, S->bond_end, S->bond_nbr, S->bond_mark, S->mark_chg);
  HIP_CHECK(hipMemcpyAsync(S->mark_chg_host, S->mark_chg, sizeof(int) * L, hipMemcpyDeviceToHost, S->stream));
  swr_kt_end(S, KT_MARK);
  HIP_CHECK(hipGetLastError());
  S->mark_launched = L;
  return SWR_OK;
}

// Finishes a pending relaxation: the flags copied at enqueue time are valid
// once the stream has passed them (call after a sync).  Converged when some
// launch changed nothing; otherwise relaxes further, synchronously.  Returns
// 1 when it gave up and raised SWR_F_MARK_DEPTH on the device.
int swr_mark_settle(swr_system *S) {
  if (S->mark_launched <= 0) return SWR_OK;
  const int n = S->n, N = S->N;
  int launch = S->mark_launched;
  S->mark_launched = 0;
  for (int k = 0; k < launch; k++)
    if (S->mark_chg_host[k] == 0) {
      S->mark_depth_hint = k * MARK_INNER;
      return SWR_OK;
    }
  while (launch < MARK_MAX_LAUNCH) {
    const int l0 = launch;
    for (int k = 0; k < 4 && launch < MARK_MAX_LAUNCH; k++, launch++)
      hipLaunchKernelGGL(k_bond_mark, dim3(div_up(N - n, 256)), dim3(256), 0, S->stream, n, N, launch, MARK_INNER,
                         S->bond_index, S->bond_end, S->bond_nbr, S->bond_mark, S->mark_chg);
    HIP_CHECK(hipMemcpyAsync(S->mark_chg_host + l0, S->mark_chg + l0, sizeof(int) * (launch - l0),
                             hipMemcpyDeviceToHost, S->stream));
    HIP_CHECK(hipStreamSynchronize(S->stream));
    for (int k = l0; k < launch; k++)
      if (S->mark_chg_host[k] == 0) {
        S->mark_depth_hint = k * MARK_INNER;
        return SWR_OK;
      }
  }
  hipLaunchKernelGGL(k_set_flag, dim3(1), dim3(1), 0, S->stream, S->status, (unsigned)SWR_F_MARK_DEPTH);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(S->stream));
  return 1;
}

// Bookkeeping after the host has synchronised with the system's stream.
// Returns 1 when it changed the device status block after that sync.
int swr_after_sync(swr_system *S) {
  if (S->status_copy_queued) {
    S->status_copy_queued = false;
    S->next_totals_known = true;
  }
  return swr_mark_settle(S) == 1;
}

// Waits for and finishes a pending relaxation before list arrays are
// reallocated or the atoms change.
int swr_quiesce(swr_system *S) {
  if (S->mark_launched <= 0) return SWR_OK;
  HIP_CHECK(hipStreamSynchronize(S->stream));
  (void)swr_after_sync(S);
  return SWR_OK;
}
