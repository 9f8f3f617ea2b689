// Time-displaced Green's functions and measurements on the device (tau.hip).
#pragma once

#include "alfq_chain.h"
#include "alfq_common.h"
#include "alfq_obs.h"

// Head of one chain's time-displaced sums: count and sign sum of the TAU_M calls
// (Obs_tau%N, Obs_tau%Ave_sign of the reference).
struct TauHead {
  double sign;
  long long n;
};

// Control_Precision_tau accumulators (NCG_tau, XMAX_tau, XMEAN_tau), per chain.
struct TauCtl {
  long long ncg;
  double xmax;
  double xmean;
  double pad;
};

// G00 = GT0 = GTT = G, G0T = G - 1 (TAU_M at tau = 0, tau_m_mod.F90:85-96).
int tau_init(const cplx* G, cplx* G00, cplx* GT0, cplx* GTT, cplx* G0T, int n, int nchain, long long cs,
             hipStream_t s);

// CGR2_2's scale-split 2n x 2n matrix H (ld 2n) and its diagonal blocks F1, F2 (ld n), from
// X = R~^-1 D^-1 Q^H of V_L P = Q D R~ (so V_L^-1 = P X, row r of V_L^-1 is row pinv[r] of
// X), the left state (UL, DL) and the right state (UR, DR, VR); *flag = 1 when DL(1) > DR(1).
int cgr22_build(const cplx* X, const int* pinv, const cplx* UL, const cplx* DL, const cplx* UR, const cplx* DR,
                const cplx* VR, cplx* H, cplx* F1, cplx* F2, int* flag, int n, int nchain, long long cs,
                hipStream_t s);

// Workgroups of the block extraction (part holds 8 doubles per workgroup and chain).
constexpr int CGR22_PARTS = 64;

// The blocks of O^-1 (2n x 2n, ld 2n) into G00, G0T, GT0, GTT, with Control_Precision_tau of
// (GR, G00), (old GTT, GTT), (old GT0, GT0), (old G0T, G0T) into ctl (partials in part).
int cgr22_extract(const cplx* Oinv, const int* flag, const cplx* GR, cplx* G00, cplx* G0T, cplx* GT0, cplx* GTT,
                  double* part, TauCtl* ctl, int n, int nchain, long long cs, hipStream_t s);

// ObserT of the Hubbard plugin at time slice nt (0..Ltrot) into the chains' blocks of `acc`
// (tau_cs bytes each: TauHead, Den0[norb] at off_den0, [3][ntau][ncell][norb][norb] at
// off_latt), with the geometry tables of the equal-time set-up.
int tau_measure(const ObsDev& o, const cplx* GT0, const cplx* G0T, const cplx* G00, const cplx* GTT,
                const ChainState* st, char* acc, long long tau_cs, long long off_den0, long long off_latt, int nt,
                int ntau, int nchain, long long cs, hipStream_t s);
