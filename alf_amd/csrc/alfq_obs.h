// Equal-time measurement buffers of the Hubbard plugin (obs.hip).
#pragma once

#include "alfq_chain.h"
#include "alfq_common.h"

// Head of one chain's measurement block: sums of the scalars and the counts and sign sums
// of the two observable groups (Obs_scal and Obs_eq of the reference).
struct ObsHead {
  cplx scal[4];             // Kin, Pot, Part, Ener
  double sign_scal, sign_eq;
  long long n_scal, n_eq;
};

// Device view of the measurement set-up: a block of `ocs` bytes per chain at buf + z * ocs
// holding the chain's parameters (ham_U at off_u, the nonzero hopping values at off_ov) and,
// from off_head on, its sums (ObsHead, Den0[norb] at off_den0, the correlation sums
// [3][ncell][norb][norb] at off_latt); the index tables are shared by the chains of a batch.
struct ObsDev {
  char* buf = nullptr;
  long long off_u = 0, off_ov = 0, off_head = 0, off_den0 = 0, off_latt = 0;
  int n = 0, ncell = 0, norb = 0, nnz = 0, n_sun = 0;
  const int* site_of = nullptr;   // [ncell][norb]: site of (cell, orbital), 0-based
  const int* jm = nullptr;        // [ncell][ncell]: the cell J with imj(I, J) = d
  const int* oi = nullptr;        // [nnz] row, column of the nonzero hopping entries
  const int* oj = nullptr;
};

int obs_measure(const ObsDev& o, const cplx* G, const ChainState* st, int nchain, long long cs, long long ocs,
                hipStream_t s);
