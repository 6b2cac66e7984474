// nw_geom.h -- element geometry shared by the host mesh producers.
#pragma once

namespace nw {
namespace geom {

// Area vectors of the 12 Hex8 sub-control surfaces, SCS order of
// HexSCS::determinant_scs (src/master_element/Hex8CVFEM.C:567-591).
void hex_scs(const double c[8][3], double a[12][3]);

}  // namespace geom
}  // namespace nw
