// Lattice.cpp -- D3Q19 / D3Q27 tables in the reference's direction order.
#include "Lattice.hh"

void Lattice::build(int q) {
  static const int dirs[27][3] = {
      {1, 0, 0},   {-1, 0, 0},  {0, 1, 0},   {0, -1, 0},  {0, 0, 1},    {0, 0, -1}, {1, 1, 0},
      {1, -1, 0},  {1, 0, 1},   {1, 0, -1},  {-1, 1, 0},  {-1, -1, 0},  {-1, 0, 1}, {-1, 0, -1},
      {0, 1, 1},   {0, 1, -1},  {0, -1, 1},  {0, -1, -1}, {1, 1, 1},    {1, 1, -1}, {1, -1, 1},
      {1, -1, -1}, {-1, 1, 1},  {-1, 1, -1}, {-1, -1, 1}, {-1, -1, -1}, {0, 0, 0}};
  stencil_.resize(q);
  inv_cs2 = 3.0;
  cs2 = 1.0 / inv_cs2;
  interactDistance = 1;
  for (int s = 0; s < q; s++) {
    const int* d = (q == 19 && s == 18) ? dirs[26] : dirs[s];
    for (int k = 0; k < 3; k++) stencil_[s].dir[k] = d[k];
    const int nz = (d[0] != 0) + (d[1] != 0) + (d[2] != 0);
    if (q == 19) stencil_[s].c = nz == 0 ? 1.0 / 3.0 : (nz == 1 ? 1.0 / 18.0 : 1.0 / 36.0);
    else stencil_[s].c = nz == 0 ? 8.0 / 27.0 : (nz == 1 ? 2.0 / 27.0 : (nz == 2 ? 1.0 / 54.0 : 1.0 / 216.0));
  }
  for (int s = 0; s < q; s++)
    for (int r = 0; r < q; r++)
      if (stencil_[r].dir[0] == -stencil_[s].dir[0] && stencil_[r].dir[1] == -stencil_[s].dir[1] &&
          stencil_[r].dir[2] == -stencil_[s].dir[2])
        stencil_[s].opp = r;
}
