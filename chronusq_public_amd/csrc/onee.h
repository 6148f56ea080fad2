// One-electron integral launcher shared between onee.hip and the host code.
#pragma once
#include <hip/hip_runtime.h>
#include "gdirect_device.h"

// Raw shell data on the device: primitive ranges prim_off[s]..prim_off[s+1] into alpha/coef,
// centres origin[3*s] (bohr).
struct OneEShellData {
  const int* prim_off;
  const double* alpha;
  const double* coef;
  const double* origin;
  // several molecules: shell -> molecule, and molecule m's nuclei [mol_atom0[m], mol_atom0[m+1])
  // (both nullptr for one molecule: every nucleus acts on every pair)
  const int* shell_mol = nullptr;
  const int* mol_atom0 = nullptr;
};

// S, T, V Cartesian blocks of the shell pairs [p_begin, p_end) of one pair class
// (0 ss, 1 ps, 2 pp, 3 ds, 4 dp, 5 dd), added into zeroed row-major ncart x ncart matrices;
// max_prim bounds the primitive count of any shell, max_atoms the nuclei acting on one pair
// (natom, or the largest molecule's atom count).
void launch_onee(int pair_class, const PairInfo* pairs, int p_begin, int p_end, const OneEShellData& sd,
                 int max_prim, const double* Z, const double* xyz, int natom, int max_atoms, const double* boys,
                 int ncart, double* S, double* T, double* V, hipStream_t stream);

// Multipole moment integrals about C[3] (host array, bohr) of one pair class: order 1 .. 3 gives
// the 3, 9 or 19 Cartesian operators x, y, z; xx, xy, xz, yy, yz, zz; xxx, xxy, xxz, xyy, xyz,
// xzz, yyy, yyz, yzz, zzz, each added into its own zeroed row-major ncart x ncart block of M.
void launch_multipole(int pair_class, const PairInfo* pairs, int p_begin, int p_end, const OneEShellData& sd,
                      int max_prim, const double* C, int order, int ncart, double* M, hipStream_t stream);
