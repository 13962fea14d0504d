// ref_grid.hpp -- the reference's partGrid reading and allocation, shared by
// the harnesses that run the reference's own code (ref_driver.cpp,
// fpi_dropin.cpp).
//
// TEST INFRASTRUCTURE ONLY.  Include after the reference's
// data_structure_mod.h (it fills the reference's globals).
#pragma once

#include <fstream>
#include <vector>

// partGrid reader with the field order of main.cpp:30-58.
static int read_grid(const char *path) {
  std::ifstream fin(path);
  if (!fin) return -1;
  int n = 0;
  fin >> n;
  if (!fin || n <= 0) return -1;
  max_points = n;
  std::vector<std::vector<int>> conn(n + 1);
  point.x.assign(n + 1, 0.0);
  point.y.assign(n + 1, 0.0);
  point.left.assign(n + 1, 0);
  point.right.assign(n + 1, 0);
  point.flag_1.assign(n + 1, 0);
  point.flag_2.assign(n + 1, 0);
  point.nbhs.assign(n + 1, 0);
  point.min_dist.assign(n + 1, 0.0);
  point.nx.assign(n + 1, 0.0);
  point.ny.assign(n + 1, 0.0);
  int width = 20;
  for (int i = 1; i <= n; i++) {
    fin >> point.x[i] >> point.y[i] >> point.left[i] >> point.right[i] >>
        point.flag_1[i] >> point.flag_2[i] >> point.min_dist[i] >>
        point.nbhs[i];
    conn[i].assign(point.nbhs[i] + 1, 0);
    for (int r = 1; r <= point.nbhs[i]; r++) fin >> conn[i][r];
    if (!fin) return -1;
    if (point.nbhs[i] > width) width = point.nbhs[i];
    if (point.flag_1[i] == 0) {
      wall_points++;
      wall_points_index.push_back(i);
    } else if (point.flag_1[i] == 1) {
      interior_points++;
      interior_points_index.push_back(i);
    } else if (point.flag_1[i] == 2) {
      outer_points++;
      outer_points_index.push_back(i);
    }
    if (point.flag_2[i] > 0) shape_points++;
  }
  point.conn.assign(n + 1, std::vector<int>(width + 1, 0));
  for (int i = 1; i <= n; i++)
    for (int r = 1; r <= point.nbhs[i]; r++) point.conn[i][r] = conn[i][r];
  return width;
}

// allocate_soln() (data_structure_mod.h:123-158) followed by what
// main.cpp:59-61 does next, with the sub-stencil lists sized for the
// cloud's widest neighbour list instead of the fixed 20
// (data_structure_mod.h:141-144).
static void allocate_soln_wide(int width) {
  allocate_soln();
  point.xpos_conn.assign(max_points + 1, std::vector<int>(width + 1, 0));
  point.xneg_conn.assign(max_points + 1, std::vector<int>(width + 1, 0));
  point.ypos_conn.assign(max_points + 1, std::vector<int>(width + 1, 0));
  point.yneg_conn.assign(max_points + 1, std::vector<int>(width + 1, 0));
  local_points = max_points;
}
