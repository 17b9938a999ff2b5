// CH4-CO2 stand-in potential energy surface.
//
// The reference's CH4-CO2 PES sources are not in its repository (.MISSING_LARGE_BLOBS lists
// PES-IDS/ai_pes_ch4_co2*.c); its plugin signature is kept: the energy is a function of the
// intermolecular distance R and of the directions of the intermolecular vector (theta1, phi1)
// and of the CO2 axis (theta2, phi2) in the frame of CH4 (PES-IDS/ai_pes_ch4_co2.h:19-29,
// src/angles_handler.cpp:357-419).  The stand-in is a sum of Lennard-Jones terms between
//   CH4 sites: C at the origin, H at r_CH (+-1, +-1, +-1)/sqrt(3) with an even number of minus
//              signs (a tetrahedron in the CH4 frame),
//   CO2 sites: C at R a, O at R a +- r_CO b (a, b: unit vectors of (theta1, phi1), (theta2, phi2)),
//   4 eps_ij ((sigma_ij / r)^12 - (sigma_ij / r)^6), sigma_ij = (sigma_i + sigma_j) / 2,
//   eps_ij = sqrt(eps_i eps_j),
// times a factor that puts the minimum over all orientations at -320 cm^-1, above the reference's
// sampling bound pesmin = -342.934 cm^-1 (example_configs/CF-F-CH4-CO2.conf:27).
#pragma once

namespace hh {
namespace ch4co2 {

constexpr double kAng = 1.8897261254578281;          // bohr per angstrom
constexpr double kKelvinHartree = 3.166811563e-6;    // Hartree per kelvin
constexpr double kRCH = 1.087 * kAng;                // C-H bond, bohr
constexpr double kRCO = 1.160 * kAng;                // C-O bond, bohr
// site types: 0 CH4 carbon, 1 CH4 hydrogen, 2 CO2 carbon, 3 CO2 oxygen (sigma in angstrom,
// epsilon in kelvin)
constexpr double kSigma[4] = {3.40, 2.65, 2.80, 3.05};
constexpr double kEpsilon[4] = {51.0, 8.6, 27.0, 79.0};
// the factor that puts the site sum's minimum at -320 cm^-1 (found by a grid search over
// R in [5, 10] and the four angles refined by coordinate descent)
constexpr double kWellScale = 1.2965035682058292;
// H positions divided by r_CH / sqrt(3)
constexpr double kH[4][3] = {{1, 1, 1}, {1, -1, -1}, {-1, 1, -1}, {-1, -1, 1}};

}  // namespace ch4co2
}  // namespace hh
