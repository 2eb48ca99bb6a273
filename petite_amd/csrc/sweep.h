// The particle range of one sweep (k_sweep_setup).  The previous transport
// handed `handed_on` particles on (only the first resume_cap of them exist:
// a reservation past the buffer kept its particle stepping); the new ones are
// the store entries from `first` to the store count (at most the capacity),
// cut so that new plus resumed stay within `bound`: a draw request region
// holds `bound` requests per process, and every particle of a sweep files at
// most one request.
#pragma once

struct SweepRange {
  int64_t first;  // first new store entry
  int64_t n;      // new entries taken
  int64_t nres;   // resumed particles
};

__host__ __device__ inline SweepRange sweep_range(int64_t first, int64_t now, int64_t capacity, int64_t bound,
                                                  int64_t handed_on, int64_t resume_cap) {
  SweepRange r;
  r.nres = handed_on > resume_cap ? resume_cap : handed_on;
  if (now > capacity) now = capacity;
  int64_t n = now - first;
  if (n > bound - r.nres) n = bound - r.nres;
  if (n < 0) n = 0;
  r.first = first;
  r.n = n;
  return r;
}
