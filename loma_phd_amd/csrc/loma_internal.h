// loma_internal.h -- host-side internal types of libloma_hip.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/loma_amd.h"

namespace loma {

// Compact-FD operators of one y grid (built by fd_build, loma_fd.cpp).
struct FdOps {
  int my = 0;
  std::vector<double> a1, b1;            // (my,7) row-major: LHS/RHS of d/deta
  std::vector<double> a2, b2;            // (my,5): LHS/RHS of d2/dy2
  std::vector<double> fmap;              // 1/(dy/deta)
  std::vector<double> inv1, up1, lo1;    // LU of a1 (p=3)
  std::vector<double> inv2, up2, lo2;    // LU of a2 (p=2)
  std::vector<double> invi, upi, loi;    // LU of b1 (integration, inty8)
};

int fd_build(int my, const double* y, FdOps& op);
void band_lu(const double* band, int n, int p, double* inv, double* up, double* lo);
void band_apply(const double* band, int n, int p, const double* u, double* out);
void band_lu_solve(const double* inv, const double* up, const double* lo, int n,
                   int p, double* b);

// wavefront-scan tables of a fixed banded factor (loma_tables.cpp): writes
// H[6][P][P][64] and returns the number of scan levels needed.  Lower solve:
// lo = multipliers (row k into row k+i at lo[k*P+i-1]); upper solve: up, inv.
int scan_table(int my, int M, int P, bool upper, const double* lo, const double* up,
               const double* inv, double* H);

void set_error(const std::string& msg);

}  // namespace loma
