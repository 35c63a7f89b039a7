// Lattice tables for device code: direction order, weights and opposites of
// the reference descriptors (modules/core/c++/descriptor/d2q9.hpp:19-53,
// d3q19.hpp:9-70, d3q27.hpp:17-98).  The order is part of the interface: the
// Zou/He boundaries index populations by number.
#pragma once
#include <cstdint>

namespace lbmx {

struct D2Q9 {
  static constexpr int D = 2, Q = 9;
  static constexpr int cx[9] = {0, -1, 1, 0, 0, -1, 1, -1, 1};
  static constexpr int cy[9] = {0, 0, 0, -1, 1, -1, 1, 1, -1};
  static constexpr int cz[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  static constexpr int opp[9] = {0, 2, 1, 4, 3, 6, 5, 8, 7};
  static constexpr double w[9] = {4. / 9., 1. / 9., 1. / 9., 1. / 9., 1. / 9.,
                                  1. / 36., 1. / 36., 1. / 36., 1. / 36.};
};

struct D3Q19 {
  static constexpr int D = 3, Q = 19;
  static constexpr int cx[19] = {0, -1, 1, 0, 0, 0, 0, -1, 1, -1, 1, -1, 1, -1, 1, 0, 0, 0, 0};
  static constexpr int cy[19] = {0, 0, 0, -1, 1, 0, 0, -1, -1, 1, 1, 0, 0, 0, 0, -1, 1, -1, 1};
  static constexpr int cz[19] = {0, 0, 0, 0, 0, -1, 1, 0, 0, 0, 0, -1, -1, 1, 1, -1, -1, 1, 1};
  static constexpr int opp[19] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};
  static constexpr double w[19] = {1. / 3.,  1. / 18., 1. / 18., 1. / 18., 1. / 18.,
                                   1. / 18., 1. / 18., 1. / 36., 1. / 36., 1. / 36.,
                                   1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36.,
                                   1. / 36., 1. / 36., 1. / 36., 1. / 36.};
};

struct D3Q27 {
  static constexpr int D = 3, Q = 27;
  static constexpr int cx[27] = {0, -1, 1, 0, -1, 1, 0, -1, 1, 0, -1, 1, 0, -1,
                                 1, 0, -1, 1, 0, -1, 1, 0, -1, 1, 0, -1, 1};
  static constexpr int cy[27] = {0, 0, 0, -1, -1, -1, 1, 1, 1, 0, 0, 0, -1, -1,
                                 -1, 1, 1, 1, 0, 0, 0, -1, -1, -1, 1, 1, 1};
  static constexpr int cz[27] = {0, 0, 0, 0, 0, 0, 0, 0, 0, -1, -1, -1, -1, -1,
                                 -1, -1, -1, -1, 1, 1, 1, 1, 1, 1, 1, 1, 1};
  static constexpr int opp[27] = {0,  2,  1,  6,  8,  7,  3,  5,  4,  18, 20, 19, 24, 26,
                                  25, 21, 23, 22, 9,  11, 10, 15, 17, 16, 12, 14, 13};
  static constexpr double w[27] = {
      8. / 27., 2. / 27.,  2. / 27.,  2. / 27., 1. / 54.,  1. / 54.,  2. / 27.,
      1. / 54., 1. / 54.,  2. / 27.,  1. / 54., 1. / 54.,  1. / 54.,  1. / 216.,
      1. / 216., 1. / 54., 1. / 216., 1. / 216., 2. / 27., 1. / 54.,  1. / 54.,
      1. / 54., 1. / 216., 1. / 216., 1. / 54., 1. / 216., 1. / 216.};
};

template <class L>
__host__ __device__ constexpr int cdir(int i, int d) {
  return d == 0 ? L::cx[i] : (d == 1 ? L::cy[i] : L::cz[i]);
}

}  // namespace lbmx
