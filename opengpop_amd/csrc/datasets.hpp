// Host-side datasets of the engine: ICGEM gravity field, IERS EOP (finals2000A), IERS 2010
// nutation tables 5.3a/b, NTP leap-second list, and the device-ready gravity records.
//
// File formats and parsing rules follow lib/force_model/src/utility/dataset_loaders.cu:30-299
// (gfc lines "gfc L M C S", EOP fixed columns 8-15/19-27/38-46/59-68 with rows shorter than
// 68 characters skipped, nutation rows "i A A'' + 14 integer multipliers" in a j = 0 and a
// j = 1 block, leap-second rows "NTP-seconds TAI-UTC").
#pragma once
#include <string>
#include <vector>

namespace ogp {

struct NutationTerm {
    double sin_uas, cos_uas; // A_i (or B''_i) and A''_i (or B_i), micro-arcseconds
    int mult[14];            // l l' F D Om L_Me L_Ve L_E L_Ma L_J L_Sa L_U L_Ne p_A
};

struct EopRecord {
    double mjd, xp, yp, ut1_utc; // MJD(UTC), pole (rad), UT1-UTC (s)
};

struct FrameData {
    std::vector<NutationTerm> lon0, lon1, obl0, obl1;
    std::vector<EopRecord> eop;
    std::vector<double> leap_mjd, leap_sec;
};

// One record per (n, m) visited by the device's column recursion, column-major (m outer,
// n = m..L inner). The column recursion Pbar_n = xi_n eps Pbar_{n-1} - eta_n Pbar_{n-2} runs on
// the rescaled functions U_n = Pbar_n / c_n, with c_m = c_{m+1} = 1 and c_n = eta_n c_{n-2}, so
// that U_n = xi'_n eps U_{n-1} - U_{n-2} with xi'_n = xi_n c_{n-1} / c_n (one multiplication
// less per degree; c_n stays within [0.46, 1.13] up to degree 70). The c_n go into the
// coefficients. Seven values per record:
//   xi'     : xi'_n (unused on the diagonal n = m)
//   c, s    : m c_n Cbar_nm, m c_n Sbar_nm                   (J and K sums, which carry m)
//   kc, ks  : c_n (1 + n + m) Cbar_nm, c_n (1 + n + m) Sbar_nm (Gamma sum)
//   zc, zs  : c_n zeta_{n,m-1} Cbar_{n,m-1}, c_n zeta_{n,m-1} Sbar_{n,m-1} (H sum carried by
//             column m)
// Coefficients of degrees n < 2 are zero: the central term is added separately
// (gravity_source_device.cu:146-152, :219).
constexpr int kRecordLen = 7;
// The gravity gradient's extension of each record (same order as the records), four values:
//   n       : the degree, weight of the radial derivative ((R/r)^n differentiates to -n/r)
//   tc, ts  : c_n zeta_{n,m-2} zeta_{n,m-1} Cbar_{n,m-2}, the same with Sbar (the second
//             derivative in sin(phi) of the term (n, m - 2), carried by column m); zero for m < 2
//   0       : padding (a record is one 32-byte scalar load)
// The other second-derivative and radial sums reuse the seven record values (see the gradient
// evaluation in device.hpp).
constexpr int kJacExtLen = 4;
struct GravityModel {
    int L = 0;         // degree/order used
    int file_max = 0;  // max_degree of the file
    double gm = 0, radius = 0;
    std::vector<double> C, S;       // triangular, index n(n+1)/2 + m
    std::vector<double> records;    // kRecordLen doubles per record
    std::vector<float> records_f;   // same in FP32
    std::vector<double> jac_ext;    // kJacExtLen doubles per record
    std::vector<double> diag_ratio; // Pbar_mm / Pbar_{m-1,m-1} (m >= 2), [1] = sqrt(3)
    size_t num_records() const { return records.size() / kRecordLen; }
};

// Throws std::runtime_error with a message naming the file on any failure.
GravityModel load_gravity(const std::string& path, int degree);
FrameData load_frames(const std::string& dir);

} // namespace ogp
