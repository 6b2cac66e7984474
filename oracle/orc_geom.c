/*
 * orc_geom.c -- oracle restatement of the Hex8 CVFEM geometry and of the
 * edge/node input producers (edge area vectors, dual volumes, edge nodal
 * gradient, edge mass flow rate).  Test infrastructure only; see orc.h.
 */
#include "orc.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* include/master_element/Hex8CVFEM.h:292-293 */
const int orc_hex_lrscv[24] = {0, 1, 1, 2, 2, 3, 0, 3, 4, 5, 5, 6,
                               6, 7, 4, 7, 0, 4, 1, 5, 2, 6, 3, 7};

/* STK Hex8 edge node ordinals (edge k joins these element-local nodes). */
static const int hex_edge_nodes[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0},
                                          {4, 5}, {5, 6}, {6, 7}, {7, 4},
                                          {0, 4}, {1, 5}, {2, 6}, {3, 7}};

/* subdivide_hex_8, include/master_element/Hex8GeometryFunctions.h:258-330:
 * the 8 corners, 12 edge midpoints, 6 face centroids and the centroid, in
 * the face-by-face order that the SCS/SCV tables below index. */
static void subdivide_hex_8(const double c[24], double v[27][3])
{
  for (int n = 0; n < 8; ++n)
    for (int d = 0; d < 3; ++d) v[n][d] = c[3 * n + d];
#define C(n, d) c[3 * (n) + (d)]
  for (int d = 0; d < 3; ++d) {
    v[8][d] = 0.5 * (C(0, d) + C(1, d));
    v[9][d] = 0.5 * (C(1, d) + C(2, d));
    v[10][d] = 0.5 * (C(2, d) + C(3, d));
    v[11][d] = 0.5 * (C(3, d) + C(0, d));
    v[12][d] = 0.25 * (C(0, d) + C(1, d) + C(2, d) + C(3, d));
    v[13][d] = 0.5 * (C(4, d) + C(5, d));
    v[14][d] = 0.5 * (C(5, d) + C(6, d));
    v[15][d] = 0.5 * (C(6, d) + C(7, d));
    v[16][d] = 0.5 * (C(7, d) + C(4, d));
    v[17][d] = 0.25 * (C(4, d) + C(5, d) + C(6, d) + C(7, d));
    v[18][d] = 0.5 * (C(1, d) + C(5, d));
    v[19][d] = 0.5 * (C(0, d) + C(4, d));
    v[20][d] = 0.25 * (C(0, d) + C(1, d) + C(4, d) + C(5, d));
    v[21][d] = 0.5 * (C(3, d) + C(7, d));
    v[22][d] = 0.5 * (C(2, d) + C(6, d));
    v[23][d] = 0.25 * (C(2, d) + C(3, d) + C(6, d) + C(7, d));
    v[24][d] = 0.25 * (C(1, d) + C(2, d) + C(5, d) + C(6, d));
    v[25][d] = 0.25 * (C(0, d) + C(3, d) + C(4, d) + C(7, d));
    double s = 0.0;
    for (int n = 0; n < 8; ++n) s += C(n, d);
    v[26][d] = s * 0.125;
  }
#undef C
}

/* quad_area_by_triangulation, Hex8GeometryFunctions.h:36-83: four
 * triangles about the quad centroid (Grandy). */
static void quad_area(const double q[4][3], double area[3])
{
  double xm[3];
  for (int d = 0; d < 3; ++d)
    xm[d] = 0.25 * (q[0][d] + q[1][d] + q[2][d] + q[3][d]);
  area[0] = area[1] = area[2] = 0.0;
  double r1[3] = {q[0][0] - xm[0], q[0][1] - xm[1], q[0][2] - xm[2]};
  for (int t = 0; t < 4; ++t) {
    const int k = (t + 1) % 4;
    const double r2[3] = {q[k][0] - xm[0], q[k][1] - xm[1], q[k][2] - xm[2]};
    area[0] += r1[1] * r2[2] - r2[1] * r1[2];
    area[1] += r1[2] * r2[0] - r2[2] * r1[0];
    area[2] += r1[0] * r2[1] - r2[0] * r1[1];
    r1[0] = r2[0];
    r1[1] = r2[1];
    r1[2] = r2[2];
  }
  area[0] *= 0.5;
  area[1] *= 0.5;
  area[2] *= 0.5;
}

/* HexSCS::determinant_scs, src/master_element/Hex8CVFEM.C:567-591 */
void orc_hex_scs_areav(const double coords[24], double areav[36])
{
  static const int facet[12][4] = {
    {20, 8, 12, 26},  {24, 9, 12, 26},  {10, 12, 26, 23}, {11, 25, 26, 12},
    {13, 20, 26, 17}, {17, 14, 24, 26}, {17, 15, 23, 26}, {16, 17, 26, 25},
    {19, 20, 26, 25}, {20, 18, 24, 26}, {22, 23, 26, 24}, {21, 25, 26, 23}};
  double v[27][3];
  subdivide_hex_8(coords, v);
  for (int ics = 0; ics < 12; ++ics) {
    double q[4][3];
    for (int i = 0; i < 4; ++i)
      for (int d = 0; d < 3; ++d) q[i][d] = v[facet[ics][i]][d];
    quad_area(q, &areav[3 * ics]);
  }
}

/* TetSCS::determinant_scs, src/master_element/Tet4CVFEM.C:522-618: edge
 * midpoints, face centroids and the centroid, then one quad per edge. */
static void tet_points(const double c[12], double v[15][3])
{
#define C(n, d) c[3 * (n) + (d)]
  for (int d = 0; d < 3; ++d) {
    for (int n = 0; n < 4; ++n) v[n][d] = C(n, d);
    v[4][d] = 0.5 * (C(0, d) + C(1, d));
    v[5][d] = 0.5 * (C(1, d) + C(2, d));
    v[6][d] = 0.5 * (C(2, d) + C(0, d));
    v[7][d] = (1.0 / 3.0) * (C(0, d) + C(1, d) + C(2, d));
    v[8][d] = 0.5 * (C(2, d) + C(3, d));
    v[9][d] = 0.5 * (C(3, d) + C(1, d));
    v[10][d] = (1.0 / 3.0) * (C(1, d) + C(2, d) + C(3, d));
    v[11][d] = 0.5 * (C(0, d) + C(3, d));
    v[12][d] = (1.0 / 3.0) * (C(0, d) + C(2, d) + C(3, d));
    v[13][d] = (1.0 / 3.0) * (C(0, d) + C(1, d) + C(3, d));
    v[14][d] = 0.0;
    for (int n = 0; n < 4; ++n) v[14][d] += 0.25 * C(n, d);
  }
#undef C
}

void orc_tet_scs_areav(const double c[12], double areav[18])
{
  static const int facet[6][4] = {{4, 7, 14, 13},  {7, 14, 10, 5},
                                  {6, 12, 14, 7},  {11, 13, 14, 12},
                                  {13, 9, 10, 14}, {10, 8, 12, 14}};
  double v[15][3];
  tet_points(c, v);
  for (int ics = 0; ics < 6; ++ics) {
    double q[4][3];
    for (int i = 0; i < 4; ++i)
      for (int d = 0; d < 3; ++d) q[i][d] = v[facet[ics][i]][d];
    quad_area(q, &areav[3 * ics]);
  }
}

/* WedSCS::determinant_scs, src/master_element/Wed6CVFEM.C:541-646 */
static void wed_points(const double c[18], double v[21][3])
{
#define C(n, d) c[3 * (n) + (d)]
  for (int d = 0; d < 3; ++d) {
    for (int n = 0; n < 6; ++n) v[n][d] = C(n, d);
    v[6][d] = 0.5 * (C(0, d) + C(1, d));
    v[7][d] = 0.5 * (C(1, d) + C(2, d));
    v[8][d] = 0.5 * (C(2, d) + C(0, d));
    v[9][d] = (1.0 / 3.0) * (C(0, d) + C(1, d) + C(2, d));
    v[10][d] = 0.5 * (C(3, d) + C(4, d));
    v[11][d] = 0.5 * (C(4, d) + C(5, d));
    v[12][d] = 0.5 * (C(5, d) + C(3, d));
    v[13][d] = (1.0 / 3.0) * (C(3, d) + C(4, d) + C(5, d));
    v[14][d] = 0.5 * (C(1, d) + C(4, d));
    v[15][d] = 0.5 * (C(0, d) + C(3, d));
    v[16][d] = 0.25 * (C(0, d) + C(1, d) + C(4, d) + C(3, d));
    v[17][d] = 0.5 * (C(2, d) + C(5, d));
    v[18][d] = 0.25 * (C(1, d) + C(4, d) + C(5, d) + C(2, d));
    v[19][d] = 0.25 * (C(5, d) + C(3, d) + C(0, d) + C(2, d));
    v[20][d] = 0.0;
    for (int n = 0; n < 6; ++n) v[20][d] += (1.0 / 6.0) * C(n, d);
  }
#undef C
}

void orc_wed_scs_areav(const double c[18], double areav[27])
{
  static const int facet[9][4] = {
    {6, 9, 20, 16},  {7, 9, 20, 18},  {9, 8, 19, 20},
    {10, 16, 20, 13}, {13, 11, 18, 20}, {12, 13, 20, 19},
    {15, 16, 20, 19}, {16, 14, 18, 20}, {19, 20, 18, 17}};
  double v[21][3];
  wed_points(c, v);
  for (int ics = 0; ics < 9; ++ics) {
    double q[4][3];
    for (int i = 0; i < 4; ++i)
      for (int d = 0; d < 3; ++d) q[i][d] = v[facet[ics][i]][d];
    quad_area(q, &areav[3 * ics]);
  }
}

/* Pyr5 points, PyrSCS::determinant_scs / PyrSCV::determinant_scv
 * (src/master_element/Pyr5CVFEM.C:437-571, 771-900): the 5 vertices, the
 * base edge midpoints (5-8), the base centroid (9), apex edge midpoints and
 * triangle centroids (10-17) and the element centroid (18), formed in the
 * reference's order of operations. */
static void pyr_points(const double c[15], double v[19][3])
{
#define C(n, d) c[3 * (n) + (d)]
  const double one3rd = 1.0 / 3.0;
  for (int d = 0; d < 3; ++d) {
    for (int n = 0; n < 5; ++n) v[n][d] = C(n, d);
    v[5][d] = 0.5 * (C(0, d) + C(1, d));
    v[6][d] = 0.5 * (C(1, d) + C(2, d));
    v[7][d] = 0.5 * (C(2, d) + C(3, d));
    v[8][d] = 0.5 * (C(3, d) + C(0, d));
    v[9][d] = 0.25 * (C(0, d) + C(1, d) + C(2, d) + C(3, d));
    v[10][d] = 0.5 * (C(1, d) + C(4, d));
    v[11][d] = 0.5 * (C(4, d) + C(0, d));
    v[12][d] = one3rd * (C(0, d) + C(1, d) + C(4, d));
    v[13][d] = 0.5 * (C(2, d) + C(4, d));
    v[14][d] = one3rd * (C(1, d) + C(2, d) + C(4, d));
    v[15][d] = 0.5 * (C(3, d) + C(4, d));
    v[16][d] = one3rd * (C(3, d) + C(4, d) + C(2, d));
    v[17][d] = one3rd * (C(0, d) + C(4, d) + C(3, d));
    v[18][d] = 0.0;
    for (int n = 0; n < 5; ++n) v[18][d] += 0.2 * C(n, d);
  }
#undef C
}

/* PyrSCS::determinant_scs (Pyr5CVFEM.C:771-900): 12 surfaces, each a quad
 * of the points above; the surfaces between a base node and the apex are
 * two triangles each (written as quads with a repeated point), the inner
 * one towards the element centroid first. */
void orc_pyr_scs_areav(const double c[15], double areav[36])
{
  static const int facet[12][4] = {
    {5, 9, 18, 12},   {6, 9, 18, 14},   {7, 9, 18, 16},   {8, 17, 18, 9},
    {12, 12, 18, 17}, {11, 12, 12, 17}, {14, 14, 18, 12}, {10, 14, 14, 12},
    {16, 16, 18, 14}, {13, 16, 16, 14}, {17, 17, 18, 16}, {15, 17, 17, 16}};
  double v[19][3];
  pyr_points(c, v);
  for (int ics = 0; ics < 12; ++ics) {
    double q[4][3];
    for (int i = 0; i < 4; ++i)
      for (int d = 0; d < 3; ++d) q[i][d] = v[facet[ics][i]][d];
    quad_area(q, &areav[3 * ics]);
  }
}

/* include/master_element/Pyr5CVFEM.h:300-305 (lrscv_): surfaces 4-11 come
 * in pairs on the four apex edges. */
const int orc_pyr_lrscv[24] = {0, 1, 1, 2, 2, 3, 0, 3, 0, 4, 0, 4,
                               1, 4, 1, 4, 2, 4, 2, 4, 3, 4, 3, 4};

/* include/master_element/Tet4CVFEM.h:266, Wed6CVFEM.h:268 */
const int orc_tet_lrscv[12] = {0, 1, 1, 2, 0, 2, 0, 3, 1, 3, 2, 3};
const int orc_wed_lrscv[18] = {0, 1, 1, 2, 0, 2, 3, 4, 4, 5,
                               3, 5, 0, 3, 1, 4, 2, 5};

/* edge of sorted (L, R) pairs by binary search, -1 if absent */
static int64_t find_edge(const int32_t* edges, int64_t nedge, int32_t a,
                         int32_t b)
{
  int64_t lo = 0, hi = nedge - 1;
  while (lo <= hi) {
    const int64_t m = (lo + hi) / 2;
    const int32_t ea = edges[2 * m], eb = edges[2 * m + 1];
    if (ea == a && eb == b) return m;
    if (ea < a || (ea == a && eb < b))
      lo = m + 1;
    else
      hi = m - 1;
  }
  return -1;
}

/* Edge area vectors of a mixed Hex8/Tet4/Wed6/Pyr5 mesh: the element loop
 * of GeometryInteriorAlg::impl_compute_edge_area_vector
 * (src/ngp_algorithms/GeometryInteriorAlg.C:163-225) with the SCS of each
 * topology; edges sorted by (node[0], node[1]) with node[0] < node[1].
 * type: 0 Hex8, 1 Tet4, 2 Wed6, 3 Pyr5; conn [nelem][8] (unused slots
 * ignored).
 * Returns -1 if an element edge is not in the list. */
int orc_edge_area_vec_mixed(int64_t nelem, const int32_t* type,
                            const int32_t* conn, const double* coords,
                            const int32_t* edges, int64_t nedge, double* areav)
{
  memset(areav, 0, sizeof(double) * 3 * (size_t)nedge);
  for (int64_t el = 0; el < nelem; ++el) {
    const int t = type[el];
    const int npe = t == 0 ? 8 : (t == 1 ? 4 : (t == 2 ? 6 : 5));
    const int nscs = t == 0 ? 12 : (t == 1 ? 6 : (t == 2 ? 9 : 12));
    const int* lr = t == 0 ? orc_hex_lrscv
                           : (t == 1 ? orc_tet_lrscv
                                     : (t == 2 ? orc_wed_lrscv : orc_pyr_lrscv));
    double c[24], a[36];
    for (int n = 0; n < npe; ++n)
      for (int d = 0; d < 3; ++d) c[3 * n + d] = coords[3 * conn[8 * el + n] + d];
    if (t == 0)
      orc_hex_scs_areav(c, a);
    else if (t == 1)
      orc_tet_scs_areav(c, a);
    else if (t == 2)
      orc_wed_scs_areav(c, a);
    else
      orc_pyr_scs_areav(c, a);
    for (int ip = 0; ip < nscs; ++ip) {
      const int32_t nl = conn[8 * el + lr[2 * ip]], nr = conn[8 * el + lr[2 * ip + 1]];
      const int64_t e = find_edge(edges, nedge, nl < nr ? nl : nr, nl < nr ? nr : nl);
      if (e < 0) return -1;
      const double sgn = (nl == edges[2 * e]) ? 1.0 : -1.0;
      for (int d = 0; d < 3; ++d) areav[3 * e + d] += a[3 * ip + d] * sgn;
    }
  }
  return 0;
}

/* hex_volume_grandy, Hex8GeometryFunctions.h:85-160 */
static double hex_volume_grandy(const double h[8][3])
{
  static const int face_nodes[6][4] = {{0, 3, 2, 1}, {4, 5, 6, 7},
                                       {0, 1, 5, 4}, {2, 3, 7, 6},
                                       {1, 2, 6, 5}, {0, 4, 3, 7}};
  static const int tri[24][3] = {
    {0, 8, 1},  {8, 2, 1},  {3, 2, 8},  {3, 8, 0},  {6, 9, 5},  {7, 9, 6},
    {4, 9, 7},  {4, 5, 9},  {10, 0, 1}, {5, 10, 1}, {4, 10, 5}, {4, 0, 10},
    {7, 6, 11}, {6, 2, 11}, {2, 3, 11}, {3, 7, 11}, {6, 12, 2}, {5, 12, 6},
    {5, 1, 12}, {1, 2, 12}, {0, 4, 13}, {4, 7, 13}, {7, 3, 13}, {3, 0, 13}};
  double cv[14][3];
  for (int n = 0; n < 8; ++n)
    for (int d = 0; d < 3; ++d) cv[n][d] = h[n][d];
  for (int k = 0; k < 6; ++k)
    for (int d = 0; d < 3; ++d)
      cv[8 + k][d] = 0.25 * (cv[face_nodes[k][0]][d] + cv[face_nodes[k][1]][d] +
                             cv[face_nodes[k][2]][d] + cv[face_nodes[k][3]][d]);
  double vol = 0.0;
  for (int k = 0; k < 24; ++k) {
    const int p = tri[k][0], q = tri[k][1], r = tri[k][2];
    const double m[3] = {cv[p][0] + cv[q][0] + cv[r][0],
                         cv[p][1] + cv[q][1] + cv[r][1],
                         cv[p][2] + cv[q][2] + cv[r][2]};
    double dx[3];
    dx[0] = (cv[q][1] - cv[p][1]) * (cv[r][2] - cv[p][2]) -
            (cv[r][1] - cv[p][1]) * (cv[q][2] - cv[p][2]);
    dx[1] = (cv[r][0] - cv[p][0]) * (cv[q][2] - cv[p][2]) -
            (cv[q][0] - cv[p][0]) * (cv[r][2] - cv[p][2]);
    dx[2] = (cv[q][0] - cv[p][0]) * (cv[r][1] - cv[p][1]) -
            (cv[r][0] - cv[p][0]) * (cv[q][1] - cv[p][1]);
    vol += m[0] * dx[0] + m[1] * dx[1] + m[2] * dx[2];
  }
  return vol / 18.0;
}

/* HexSCV::determinant_scv, src/master_element/Hex8CVFEM.C:365-389 */
void orc_hex_scv_volume(const double coords[24], double vol[8])
{
  static const int sub[8][8] = {
    {0, 8, 12, 11, 19, 20, 26, 25},  {8, 1, 9, 12, 20, 18, 24, 26},
    {12, 9, 2, 10, 26, 24, 22, 23},  {11, 12, 10, 3, 25, 26, 23, 21},
    {19, 20, 26, 25, 4, 13, 17, 16}, {20, 18, 24, 26, 13, 5, 14, 17},
    {26, 24, 22, 23, 17, 14, 6, 15}, {25, 26, 23, 21, 16, 17, 15, 7}};
  double v[27][3];
  subdivide_hex_8(coords, v);
  for (int ip = 0; ip < 8; ++ip) {
    double h[8][3];
    for (int n = 0; n < 8; ++n)
      for (int d = 0; d < 3; ++d) h[n][d] = v[sub[ip][n]][d];
    vol[ip] = hex_volume_grandy(h);
  }
}

/* TetSCV::determinant_scv, src/master_element/Tet4CVFEM.C:243-343: each
 * sub-control volume is the hexahedron of eight points of the subdivided
 * tetrahedron, its volume hex_volume_grandy's. */
void orc_tet_scv_volume(const double c[12], double vol[4])
{
  static const int sub[4][8] = {{0, 4, 7, 6, 11, 13, 14, 12},
                                {1, 5, 7, 4, 9, 10, 14, 13},
                                {2, 6, 7, 5, 8, 12, 14, 10},
                                {3, 9, 13, 11, 8, 10, 14, 12}};
  double v[15][3];
  tet_points(c, v);
  for (int s = 0; s < 4; ++s) {
    double h[8][3];
    for (int n = 0; n < 8; ++n)
      for (int d = 0; d < 3; ++d) h[n][d] = v[sub[s][n]][d];
    vol[s] = hex_volume_grandy(h);
  }
}

/* WedSCV::determinant_scv, src/master_element/Wed6CVFEM.C:267-369 */
void orc_wed_scv_volume(const double c[18], double vol[6])
{
  static const int sub[6][8] = {
    {0, 15, 16, 6, 8, 19, 20, 9},    {9, 6, 1, 7, 20, 16, 14, 18},
    {8, 9, 7, 2, 19, 20, 18, 17},    {19, 15, 16, 20, 12, 3, 10, 13},
    {20, 16, 14, 18, 13, 10, 4, 11}, {19, 20, 18, 17, 12, 13, 11, 5}};
  double v[21][3];
  wed_points(c, v);
  for (int s = 0; s < 6; ++s) {
    double h[8][3];
    for (int n = 0; n < 8; ++n)
      for (int d = 0; d < 3; ++d) h[n][d] = v[sub[s][n]][d];
    vol[s] = hex_volume_grandy(h);
  }
}

/* bhex_volume_grandy (Hex8GeometryFunctions.h:163-254): Grandy's volume
 * of a hexahedron whose top face (4, 5, 6, 7) is two triangles split along
 * 5-7 (its "centroid" is the midpoint of that diagonal). */
static double bhex_volume_grandy(const double h[8][3])
{
  static const int tri[24][3] = {
    {0, 8, 1},  {8, 2, 1},  {3, 2, 8},  {3, 8, 0},  {6, 9, 5},  {7, 9, 6},
    {4, 9, 7},  {4, 5, 9},  {10, 0, 1}, {5, 10, 1}, {4, 10, 5}, {4, 0, 10},
    {7, 6, 11}, {6, 2, 11}, {2, 3, 11}, {3, 7, 11}, {6, 12, 2}, {5, 12, 6},
    {5, 1, 12}, {1, 2, 12}, {0, 4, 13}, {4, 7, 13}, {7, 3, 13}, {3, 0, 13}};
  double cv[14][3];
  for (int n = 0; n < 8; ++n)
    for (int d = 0; d < 3; ++d) cv[n][d] = h[n][d];
  for (int d = 0; d < 3; ++d) {
    cv[8][d] = 0.25 * (h[0][d] + h[1][d] + h[2][d] + h[3][d]);
    cv[9][d] = 0.5 * (h[5][d] + h[7][d]);
    cv[10][d] = 0.25 * (h[0][d] + h[1][d] + h[5][d] + h[4][d]);
    cv[11][d] = 0.25 * (h[3][d] + h[2][d] + h[6][d] + h[7][d]);
    cv[12][d] = 0.25 * (h[1][d] + h[2][d] + h[6][d] + h[5][d]);
    cv[13][d] = 0.25 * (h[0][d] + h[3][d] + h[7][d] + h[4][d]);
  }
  double vol = 0.0;
  for (int k = 0; k < 24; ++k) {
    const int p = tri[k][0], q = tri[k][1], r = tri[k][2];
    const double m[3] = {cv[p][0] + cv[q][0] + cv[r][0],
                         cv[p][1] + cv[q][1] + cv[r][1],
                         cv[p][2] + cv[q][2] + cv[r][2]};
    double dx[3];
    dx[0] = (cv[q][1] - cv[p][1]) * (cv[r][2] - cv[p][2]) -
            (cv[r][1] - cv[p][1]) * (cv[q][2] - cv[p][2]);
    dx[1] = (cv[r][0] - cv[p][0]) * (cv[q][2] - cv[p][2]) -
            (cv[q][0] - cv[p][0]) * (cv[r][2] - cv[p][2]);
    dx[2] = (cv[q][0] - cv[p][0]) * (cv[r][1] - cv[p][1]) -
            (cv[r][0] - cv[p][0]) * (cv[q][1] - cv[p][1]);
    vol += m[0] * dx[0] + m[1] * dx[1] + m[2] * dx[2];
  }
  return vol / 18.0;
}

/* octohedron_volume_by_triangle_facets + polyhedral_volume_by_faces
 * (Pyr5CVFEM.C:349-433): the apex sub-volume from its 10 vertices, the four
 * non-planar faces split at the midpoints of (3, 9), (3, 5), (5, 7), (7, 9),
 * as 24 triangles and the divergence theorem. */
static double octahedron_volume(const double v10[10][3])
{
  static const int tri[24][3] = {
    {1, 3, 10}, {2, 10, 3}, {2, 9, 10}, {10, 9, 1}, {4, 3, 11}, {3, 1, 11},
    {11, 1, 5}, {4, 11, 5}, {1, 12, 5}, {1, 7, 12}, {12, 7, 6}, {5, 12, 6},
    {9, 8, 13}, {13, 8, 7}, {13, 7, 1}, {9, 13, 1}, {4, 5, 0},  {5, 6, 0},
    {6, 7, 0},  {7, 8, 0},  {0, 8, 9},  {0, 9, 2},  {0, 2, 3},  {0, 3, 4}};
  double cv[14][3];
  for (int n = 0; n < 10; ++n)
    for (int d = 0; d < 3; ++d) cv[n][d] = v10[n][d];
  for (int d = 0; d < 3; ++d) {
    cv[10][d] = 0.5 * (v10[3][d] + v10[9][d]);
    cv[11][d] = 0.5 * (v10[3][d] + v10[5][d]);
    cv[12][d] = 0.5 * (v10[5][d] + v10[7][d]);
    cv[13][d] = 0.5 * (v10[7][d] + v10[9][d]);
  }
  double vol = 0.0;
  for (int k = 0; k < 24; ++k) {
    const int p = tri[k][0], q = tri[k][1], r = tri[k][2];
    double xf[3];
    for (int d = 0; d < 3; ++d) xf[d] = cv[p][d] + cv[q][d] + cv[r][d];
    vol = vol +
          xf[0] * ((cv[q][1] - cv[p][1]) * (cv[r][2] - cv[p][2]) -
                   (cv[r][1] - cv[p][1]) * (cv[q][2] - cv[p][2])) -
          xf[1] * ((cv[q][0] - cv[p][0]) * (cv[r][2] - cv[p][2]) -
                   (cv[r][0] - cv[p][0]) * (cv[q][2] - cv[p][2])) +
          xf[2] * ((cv[q][0] - cv[p][0]) * (cv[r][1] - cv[p][1]) -
                   (cv[r][0] - cv[p][0]) * (cv[q][1] - cv[p][1]));
  }
  return vol / 18.0;
}

/* PyrSCV::determinant_scv (Pyr5CVFEM.C:437-571): one bent hexahedron per
 * base node, the octahedron about the apex. */
void orc_pyr_scv_volume(const double c[15], double vol[5])
{
  static const int sub[5][10] = {{0, 5, 9, 8, 11, 12, 18, 17, -1, -1},
                                 {1, 6, 9, 5, 10, 14, 18, 12, -1, -1},
                                 {2, 7, 9, 6, 13, 16, 18, 14, -1, -1},
                                 {3, 8, 9, 7, 15, 17, 18, 16, -1, -1},
                                 {4, 18, 15, 17, 11, 12, 10, 14, 13, 16}};
  double v[19][3];
  pyr_points(c, v);
  for (int s = 0; s < 4; ++s) {
    double h[8][3];
    for (int n = 0; n < 8; ++n)
      for (int d = 0; d < 3; ++d) h[n][d] = v[sub[s][n]][d];
    vol[s] = bhex_volume_grandy(h);
  }
  double o[10][3];
  for (int n = 0; n < 10; ++n)
    for (int d = 0; d < 3; ++d) o[n][d] = v[sub[4][n]][d];
  vol[4] = octahedron_volume(o);
}

/* Dual nodal volumes of a mixed Hex8/Tet4/Wed6/Pyr5 mesh: the element loop
 * of GeometryInteriorAlg::impl_compute_dual_nodal_volume
 * (src/ngp_algorithms/GeometryInteriorAlg.C:70-112); sub-volume ip of an
 * element belongs to its local node ip (ipNodeMap is the identity for all
 * four, include/master_element/{Hex8,Tet4,Wed6,Pyr5}CVFEM.h). */
void orc_dual_volume_mixed(int64_t nelem, const int32_t* type,
                           const int32_t* conn, const double* coords,
                           int64_t nnode, double* dualvol)
{
  memset(dualvol, 0, sizeof(double) * (size_t)nnode);
  for (int64_t el = 0; el < nelem; ++el) {
    const int t = type[el];
    const int npe = t == 0 ? 8 : (t == 1 ? 4 : (t == 2 ? 6 : 5));
    double c[24], v[8];
    for (int n = 0; n < npe; ++n)
      for (int d = 0; d < 3; ++d) c[3 * n + d] = coords[3 * conn[8 * el + n] + d];
    if (t == 0)
      orc_hex_scv_volume(c, v);
    else if (t == 1)
      orc_tet_scv_volume(c, v);
    else if (t == 2)
      orc_wed_scv_volume(c, v);
    else
      orc_pyr_scv_volume(c, v);
    for (int ip = 0; ip < npe; ++ip) dualvol[conn[8 * el + ip]] += v[ip];
  }
}

int64_t orc_create_edges_hex(int64_t nelem, const int32_t* conn,
                             int64_t nnode, int32_t* edges, int64_t edges_cap,
                             int32_t* elem_edges)
{
  /* per-node singly linked list of edges whose node[0] is that node */
  int64_t* head = (int64_t*)malloc(sizeof(int64_t) * (size_t)nnode);
  int64_t* next = (int64_t*)malloc(sizeof(int64_t) * (size_t)edges_cap);
  for (int64_t i = 0; i < nnode; ++i) head[i] = -1;
  int64_t ne = 0;
  for (int64_t el = 0; el < nelem; ++el) {
    for (int k = 0; k < 12; ++k) {
      int32_t a = conn[8 * el + hex_edge_nodes[k][0]];
      int32_t b = conn[8 * el + hex_edge_nodes[k][1]];
      if (b < a) {
        int32_t t = a;
        a = b;
        b = t;
      }
      int64_t found = -1;
      for (int64_t it = head[a]; it >= 0; it = next[it])
        if (edges[2 * it + 1] == b) {
          found = it;
          break;
        }
      if (found < 0) {
        if (ne >= edges_cap) {
          free(head);
          free(next);
          return -1;
        }
        edges[2 * ne] = a;
        edges[2 * ne + 1] = b;
        next[ne] = head[a];
        head[a] = ne;
        found = ne++;
      }
      elem_edges[12 * el + k] = (int32_t)found;
    }
  }
  free(head);
  free(next);
  return ne;
}

void orc_edge_area_vec_hex(int64_t nelem, const int32_t* conn,
                           const double* coords, const int32_t* elem_edges,
                           const int32_t* edges, int64_t nedge, double* areav)
{
  memset(areav, 0, sizeof(double) * 3 * (size_t)nedge);
  for (int64_t el = 0; el < nelem; ++el) {
    double c[24], a[36];
    for (int n = 0; n < 8; ++n)
      for (int d = 0; d < 3; ++d) c[3 * n + d] = coords[3 * conn[8 * el + n] + d];
    orc_hex_scs_areav(c, a);
    for (int ip = 0; ip < 12; ++ip) {
      const int32_t e = elem_edges[12 * el + ip]; /* scsIpEdgeOrd = identity */
      const int32_t lnElem = conn[8 * el + orc_hex_lrscv[2 * ip]];
      const double sgn = (lnElem == edges[2 * e]) ? 1.0 : -1.0;
      for (int d = 0; d < 3; ++d) areav[3 * e + d] += a[3 * ip + d] * sgn;
    }
  }
}

void orc_dual_volume_hex(int64_t nelem, const int32_t* conn,
                         const double* coords, int64_t nnode, double* dualvol)
{
  memset(dualvol, 0, sizeof(double) * (size_t)nnode);
  for (int64_t el = 0; el < nelem; ++el) {
    double c[24], v[8];
    for (int n = 0; n < 8; ++n)
      for (int d = 0; d < 3; ++d) c[3 * n + d] = coords[3 * conn[8 * el + n] + d];
    orc_hex_scv_volume(c, v);
    /* HexSCV ipNodeMap is the identity: sub-volume ip belongs to node ip */
    for (int ip = 0; ip < 8; ++ip) dualvol[conn[8 * el + ip]] += v[ip];
  }
}

void orc_nodal_grad_edge(int64_t nnode, int64_t nedge, const int32_t* edges,
                         const double* areav, const double* dualvol, int dim1,
                         const double* phi, double* grad)
{
  memset(grad, 0, sizeof(double) * (size_t)nnode * (size_t)dim1 * 3);
  for (int64_t e = 0; e < nedge; ++e) {
    const int32_t L = edges[2 * e], R = edges[2 * e + 1];
    const double invL = 1.0 / dualvol[L];
    const double invR = 1.0 / dualvol[R];
    int counter = 0;
    for (int i = 0; i < dim1; ++i) {
      const double phiIp = 0.5 * (phi[(int64_t)dim1 * L + i] +
                                  phi[(int64_t)dim1 * R + i]);
      for (int j = 0; j < 3; ++j) {
        const double a = areav[3 * e + j] * phiIp;
        grad[(int64_t)dim1 * 3 * L + counter] += a * invL;
        grad[(int64_t)dim1 * 3 * R + counter] -= a * invR;
        counter++;
      }
    }
  }
}

void orc_mdot_edge(int64_t nedge, const int32_t* edges, const double* areav,
                   const double* rho, const double* vel, double* mdot)
{
  for (int64_t e = 0; e < nedge; ++e) {
    const int32_t L = edges[2 * e], R = edges[2 * e + 1];
    double t = 0.0;
    for (int d = 0; d < 3; ++d)
      t += 0.5 * (rho[L] * vel[3 * L + d] + rho[R] * vel[3 * R + d]) *
           areav[3 * e + d];
    mdot[e] = t;
  }
}

/* MdotEdgeAlg::execute interior edges (src/ngp_algorithms/MdotEdgeAlg.C:
 * 121-186) without the balanced-buoyancy and mesh-motion terms: node
 * arrays [n][3] / [n], areav [e][3]. */
void orc_mdot_edge_alg(int64_t nedge, const int32_t* edges,
                       const double* coords, const double* vel,
                       const double* dpdx, const double* rho, const double* p,
                       const double* udiag, const double* areav, double noc,
                       double interp, const orc_buoyancy* buoy, double* mdot)
{
  for (int64_t e = 0; e < nedge; ++e) {
    const int32_t L = edges[2 * e], R = edges[2 * e + 1];
    const double* av = areav + 3 * e;
    const double projTimeScale = 0.5 * (1.0 / udiag[L] + 1.0 / udiag[R]);
    const double rhoIp = 0.5 * (rho[L] + rho[R]);
    double axdx = 0.0, asq = 0.0, dx[3];
    for (int d = 0; d < 3; ++d) {
      dx[d] = coords[3 * R + d] - coords[3 * L + d];
      asq += av[d] * av[d];
      axdx += av[d] * dx[d];
    }
    const double inv_axdx = 1.0 / axdx;
    double t = -projTimeScale * (p[R] - p[L]) * asq * inv_axdx;
    if (buoy) { /* MdotEdgeAlg.C:147-154 */
      const double masked_weights = 0.5 * (buoy->mask[L] + buoy->mask[R]);
      for (int d = 0; d < 3; ++d)
        t += projTimeScale * av[d] * buoy->gravity[d] * rhoIp * masked_weights;
    }
    for (int d = 0; d < 3; ++d) {
      const double kxj = av[d] - asq * inv_axdx * dx[d];
      const double rhoUjIp = 0.5 * (rho[R] * vel[3 * R + d] + rho[L] * vel[3 * L + d]);
      const double ujIp = 0.5 * (vel[3 * R + d] + vel[3 * L + d]);
      double GjIp = 0.5 * (dpdx[3 * R + d] / udiag[R] + dpdx[3 * L + d] / udiag[L]);
      if (buoy) /* MdotEdgeAlg.C:163-169 */
        GjIp -= 0.5 * ((buoy->mask[R] * buoy->source[3 * R + d]) / udiag[R] +
                       (buoy->mask[L] * buoy->source[3 * L + d]) / udiag[L]);
      t += (interp * rhoUjIp + (1.0 - interp) * rhoIp * ujIp + GjIp) * av[d] -
           kxj * GjIp * noc;
    }
    mdot[e] = t;
  }
}
