// Driver for the reference's own hierarchy enumeration (src/hrchy_space.cc),
// compiled from the reference sources in place into oracle/_ref/ by
// oracle/Makefile.  TEST INFRASTRUCTURE ONLY: used by tests to pin the
// oracle's breadth-first ADO order and +-1 neighbour tables.
#include <cstring>
#include "hrchy_space.h"

extern "C" int ref_hrchy(int n_dim, int max_depth, int* n_out, int* p1_out, int* m1_out,
                         int capacity) {
  libheom::hrchy_space hs;
  hs.n_dim = n_dim;
  int n_hrchy = libheom::alloc_hrchy_space(
      hs, max_depth, [](int, int) {}, 1, [](libheom::vector<int>, int) { return false; }, false);
  if (n_out == nullptr || capacity < n_hrchy) return n_hrchy;
  for (int l = 0; l < n_hrchy; ++l)
    for (int k = 0; k < n_dim; ++k) n_out[l * n_dim + k] = hs.n[l][k];
  for (int l = 0; l <= n_hrchy; ++l)
    for (int k = 0; k < n_dim; ++k) {
      p1_out[l * n_dim + k] = hs.ptr_p1[l][k];
      m1_out[l * n_dim + k] = hs.ptr_m1[l][k];
    }
  return n_hrchy;
}

// the same enumeration with an hrchy_filter (hrchy_space.h:29-36)
extern "C" int ref_hrchy_filtered(int n_dim, int max_depth,
                                  int (*filter)(const int*, int, int, void*), void* user,
                                  int* n_out, int* p1_out, int* m1_out, int capacity) {
  libheom::hrchy_space hs;
  hs.n_dim = n_dim;
  int n_hrchy = libheom::alloc_hrchy_space(
      hs, max_depth, [](int, int) {}, 1,
      [&](libheom::vector<int> index, int depth) {
        return filter(index.data(), n_dim, depth, user) != 0;
      },
      true);
  if (n_out == nullptr || capacity < n_hrchy) return n_hrchy;
  for (int l = 0; l < n_hrchy; ++l)
    for (int k = 0; k < n_dim; ++k) n_out[l * n_dim + k] = hs.n[l][k];
  for (int l = 0; l <= n_hrchy; ++l)
    for (int k = 0; k < n_dim; ++k) {
      p1_out[l * n_dim + k] = hs.ptr_p1[l][k];
      m1_out[l * n_dim + k] = hs.ptr_m1[l][k];
    }
  return n_hrchy;
}
