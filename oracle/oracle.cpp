// FP64 host oracle: TEST INFRASTRUCTURE ONLY (see oracle.hpp for the scope and the rule that
// only tests/, the smoke check and bench.py's cpu_baseline leg may load it).
//
// Built with -ffp-contract=off so that every expression rounds exactly as written, in the
// reference's own left-to-right operation order.
#include "oracle.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// lib/force_model/include/force_model/constants.cuh:23-35
constexpr double kPi = 3.14159265358979324;
constexpr double kArcsPerRad = 3600.0 * 180.0 / kPi;
constexpr double kMjdJ2000 = 51544.5;
constexpr double kMjd1900 = 15020.0;
constexpr double kTtTai = 32.184;
constexpr double kRk8Mid[10] = { 0.0, 4.0 / 27.0, 2.0 / 9.0, 1.0 / 3.0, 1.0 / 2.0, 2.0 / 3.0,
    1.0 / 6.0, 1.0, 5.0 / 6.0, 1.0 };

std::string g_err;

// ---------------------------------------------------------------------------------------
// Integrator coefficient tables: gpu_constants.cuh:22-98 (Berry & Healy 2004 ordinate form)
// ---------------------------------------------------------------------------------------
const double kRk8b[10] = { 41.0 / 840.0, 0 / 840.0, 0 / 840.0, 27.0 / 840.0, 272.0 / 840.0,
    27.0 / 840.0, 216.0 / 840.0, 0 / 840.0, 216.0 / 840.0, 41.0 / 840.0 };

void rk8a(int quirks, double a[9][9])
{
    const double t[9][9] = { { 4.0 / 27, 0, 0, 0, 0, 0, 0, 0, 0 },
        { 1.0 / 18, 3.0 / 18, 0, 0, 0, 0, 0, 0, 0 },
        { 1.0 / 12, 0.0 / 12, 3.0 / 12, 0, 0, 0, 0, 0, 0 },
        { 1.0 / 8, 0, 0, 3.0 / 8, 0, 0, 0, 0, 0 },
        { 13.0 / 54, 0 / 54, -27.0 / 54, 42.0 / 54, 8.0 / 54, 0, 0, 0, 0 },
        { 389.0 / 4320, 0.0 / 4320, -54.0 / 4320, 966.0 / 4320, -824.0 / 4320.0, 243.0 / 4320,
            0, 0, 0 },
        { -234.0 / 20, 0, 81.0 / 20, -1164.0 / 20, 656.0 / 20, -122.0 / 20, 800.0 / 20, 0, 0 },
        { -127.0 / 288, 0.0 / 288, 18.0 / 288, -678.0 / 288, 456.0 / 288, -9.0 / 288,
            576.0 / 288, 4.0 / 288, 0 },
        { 1481.0 / 820, 0.0 / 820, -81.0 / 820, 7104.0 / 820, -3376.0 / 820, 72.0 / 820,
            -5040.0 / 820, -60.0 / 820, 720.0 / 820 } };
    std::memcpy(a, t, sizeof(t));
    // Row 7 of the reference sums to 17/20 while c7 = 1; -231 restores consistency (and the
    // 8th-order linear conditions). Bit 0 of quirks keeps the reference value.
    if (!(quirks & 1))
        a[6][0] = -231.0 / 20;
}

const double kGJ[10][9] = {
    { 3250433.0 / 53222400.0, 572741.0 / 5702400.0, -8701681.0 / 39916800.0,
        4026311.0 / 13305600.0, -917039.0 / 3193344.0, 7370669.0 / 39916800.0,
        -1025779.0 / 13305600.0, 754331.0 / 39916800.0, -330157.0 / 159667200.0 },
    { -330157.0 / 159667200.0, 530113.0 / 6652800.0, 518887.0 / 19958400.0, -27631.0 / 623700.0,
        44773.0 / 1064448.0, -531521.0 / 19958400.0, 109343.0 / 9979200.0, -1261.0 / 475200.0,
        45911.0 / 159667200.0 },
    { 45911.0 / 159667200.0, -185839.0 / 39916800.0, 171137.0 / 1900800.0,
        73643.0 / 39916800.0, -25775.0 / 3193344.0, 77597.0 / 13305600.0,
        -98911.0 / 39916800.0, 24173.0 / 39916800.0, -3499.0 / 53222400.0 },
    { -3499.0 / 53222400.0, 4387.0 / 4989600.0, -35039.0 / 4989600.0, 90817.0 / 950400.0,
        -20561.0 / 3193344.0, 2117.0 / 9979200.0, 2059.0 / 6652800.0, -317.0 / 2851200.0,
        317.0 / 22809600.0 },
    { 317.0 / 22809600.0, -2539.0 / 13305600.0, 55067.0 / 39916800.0, -326911.0 / 39916800.0,
        14797.0 / 152064.0, -326911.0 / 39916800.0, 55067.0 / 39916800.0,
        -2539.0 / 13305600.0, 317.0 / 22809600.0 },
    { 317.0 / 22809600.0, -317.0 / 2851200.0, 2059.0 / 6652800.0, 2117.0 / 9979200.0,
        -20561.0 / 3193344.0, 90817.0 / 950400.0, -35039.0 / 4989600.0, 4387.0 / 4989600.0,
        -3499.0 / 53222400.0 },
    { -3499.0 / 53222400.0, 24173.0 / 39916800.0, -98911.0 / 39916800.0,
        77597.0 / 13305600.0, -25775.0 / 3193344.0, 73643.0 / 39916800.0,
        171137.0 / 1900800.0, -185839.0 / 39916800.0, 45911.0 / 159667200.0 },
    { 45911.0 / 159667200.0, -1261.0 / 475200.0, 109343.0 / 9979200.0,
        -531521.0 / 19958400.0, 44773.0 / 1064448.0, -27631.0 / 623700.0,
        518887.0 / 19958400.0, 530113.0 / 6652800.0, -330157.0 / 159667200.0 },
    { -330157.0 / 159667200.0, 754331.0 / 39916800.0, -1025779.0 / 13305600.0,
        7370669.0 / 39916800.0, -917039.0 / 3193344.0, 4026311.0 / 13305600.0,
        -8701681.0 / 39916800.0, 572741.0 / 5702400.0, 3250433.0 / 53222400.0 },
    { 3250433.0 / 53222400.0, -11011481.0 / 19958400.0, 6322573.0 / 2851200.0,
        -8660609.0 / 1663200.0, 25162927.0 / 3193344.0, -159314453.0 / 19958400.0,
        18071351.0 / 3326400.0, -24115843.0 / 9979200.0, 103798439.0 / 159667200.0 }
};

const double kSA[10][9] = {
    { 19087.0 / 89600.0, -427487.0 / 725760.0, 3498217.0 / 3628800.0, -500327.0 / 403200.0,
        6467.0 / 5670.0, -2616161.0 / 3628800.0, 24019.0 / 80640.0, -263077.0 / 3628800.0,
        8183.0 / 1036800.0 },
    { 8183.0 / 1036800.0, 57251.0 / 403200.0, -1106377.0 / 3628800.0, 218483.0 / 725760.0,
        -69.0 / 280.0, 530177.0 / 3628800.0, -210359.0 / 3628800.0, 5533.0 / 403200.0,
        -425.0 / 290304.0 },
    { -425.0 / 290304.0, 76453.0 / 3628800.0, 5143.0 / 57600.0, -660127.0 / 3628800.0,
        661.0 / 5670.0, -4997.0 / 80640.0, 83927.0 / 3628800.0, -19109.0 / 3628800.0,
        7.0 / 12800.0 },
    { 7.0 / 12800.0, -23173.0 / 3628800.0, 29579.0 / 725760.0, 2497.0 / 57600.0,
        -2563.0 / 22680.0, 172993.0 / 3628800.0, -6463.0 / 403200.0, 2497.0 / 725760.0,
        -2497.0 / 7257600.0 },
    { -2497.0 / 7257600.0, 1469.0 / 403200.0, -68119.0 / 3628800.0, 252769.0 / 3628800.0, 0.0,
        -252769.0 / 3628800.0, 68119.0 / 3628800.0, -1469.0 / 403200.0, 2497.0 / 7257600.0 },
    { 2497.0 / 7257600.0, -2497.0 / 725760.0, 6463.0 / 403200.0, -172993.0 / 3628800.0,
        2563.0 / 22680.0, -2497.0 / 57600.0, -29579.0 / 725760.0, 23173.0 / 3628800.0,
        -7.0 / 12800.0 },
    { -7.0 / 12800.0, 19109.0 / 3628800.0, -83927.0 / 3628800.0, 4997.0 / 80640.0,
        -661.0 / 5670.0, 660127.0 / 3628800.0, -5143.0 / 57600.0, -76453.0 / 3628800.0,
        425.0 / 290304.0 },
    { 425.0 / 290304.0, -5533.0 / 403200.0, 210359.0 / 3628800.0, -530177.0 / 3628800.0,
        69.0 / 280.0, -218483.0 / 725760.0, 1106377.0 / 3628800.0, -57251.0 / 403200.0,
        -8183.0 / 1036800.0 },
    { -8183.0 / 1036800.0, 263077.0 / 3628800.0, -24019.0 / 80640.0, 2616161.0 / 3628800.0,
        -6467.0 / 5670.0, 500327.0 / 403200.0, -3498217.0 / 3628800.0, 427487.0 / 725760.0,
        -19087.0 / 89600.0 },
    { 25713.0 / 89600.0, -9401029.0 / 3628800.0, 5393233.0 / 518400.0, -9839609.0 / 403200.0,
        167287.0 / 4536.0, -135352319.0 / 3628800.0, 10219841.0 / 403200.0,
        -40987771.0 / 3628800.0, 3288521.0 / 1036800.0 }
};

// ---------------------------------------------------------------------------------------
// Datasets: dataset_loaders.cu:30-299, types.cuh:23-194
// ---------------------------------------------------------------------------------------
struct NutTerm {
    double s, c;
    int arg[14];
};
struct NutTables {
    std::vector<NutTerm> j0, j1;
};
struct EopRec {
    double mjd, xp, yp, dut1;
};

struct Data {
    // gravity
    unsigned N = 0;
    double GM = 0, R = 0;
    std::vector<double> C, S, xi, eta, zeta, delta, pdiag;
    // frames
    NutTables lon, obl;
    std::vector<EopRec> eop;
    std::vector<double> ls_mjd, ls_val;
};
std::unique_ptr<Data> g_data;

size_t tri(size_t r, size_t c) { return r * (r + 1) / 2 + c; }

NutTables load_nutation(const std::string& path)
{
    std::ifstream f(path);
    if (!f.is_open())
        throw std::runtime_error("cannot open " + path);
    NutTables t;
    std::string line;
    int j = 0;
    while (std::getline(f, line)) {
        if (line.find("j = 0") != std::string::npos) {
            j = 0;
            continue;
        }
        if (line.find("j = 1") != std::string::npos) {
            j = 1;
            continue;
        }
        std::istringstream is(line);
        int i;
        if (!(is >> i))
            continue;
        NutTerm term;
        if (!(is >> term.s >> term.c))
            continue;
        bool ok = true;
        for (int k = 0; k < 14; ++k)
            if (!(is >> term.arg[k])) {
                ok = false;
                break;
            }
        if (ok)
            (j == 0 ? t.j0 : t.j1).push_back(term);
    }
    return t;
}

std::vector<EopRec> load_eop(const std::string& path)
{
    std::ifstream f(path);
    if (!f.is_open())
        throw std::runtime_error("cannot open " + path);
    std::vector<EopRec> out;
    std::string line;
    const char* ws = " \t\n\r\f\v";
    while (std::getline(f, line)) {
        size_t e = line.find_last_not_of(ws);
        if (e == std::string::npos)
            line.clear();
        else
            line.erase(e + 1);
        if (line.empty() || line[0] == '#' || line.length() < 68)
            continue;
        try {
            EopRec r;
            r.mjd = std::stod(line.substr(7, 8));
            r.xp = std::stod(line.substr(18, 9)) / kArcsPerRad;
            r.yp = std::stod(line.substr(37, 9)) / kArcsPerRad;
            r.dut1 = std::stod(line.substr(58, 10));
            out.push_back(r);
        } catch (const std::exception&) {
            continue;
        }
    }
    return out;
}

void load_leap(const std::string& path, std::vector<double>& mjd, std::vector<double>& val)
{
    std::ifstream f(path);
    if (!f.is_open())
        throw std::runtime_error("cannot open " + path);
    std::string line;
    while (std::getline(f, line)) {
        size_t p = line.find_first_not_of(" \t");
        if (line.empty() || p == std::string::npos || line[p] == '#')
            continue;
        std::stringstream ss(line);
        long long ntp;
        double d;
        if (ss >> ntp >> d) {
            mjd.push_back(kMjd1900 + (static_cast<double>(ntp) / 86400.0));
            val.push_back(d);
        }
    }
}

void load_gravity(const std::string& path, int degree, Data& d)
{
    std::ifstream f(path);
    if (!f.is_open())
        throw std::runtime_error("cannot open " + path);
    std::string line;
    bool end = false;
    unsigned maxdeg = 0;
    while (std::getline(f, line)) {
        if (line.find("end_of_head") != std::string::npos) {
            end = true;
            break;
        }
        std::istringstream is(line);
        std::string tok;
        is >> tok;
        if (tok == "earth_gravity_constant")
            is >> d.GM;
        else if (tok == "radius")
            is >> d.R;
        else if (tok == "max_degree")
            is >> maxdeg;
    }
    if (!end)
        throw std::runtime_error("Invalid ICGEM format: 'end_of_head' not found.");
    if (degree >= 0 && (unsigned)degree > maxdeg)
        throw std::runtime_error("requested degree exceeds the file's max_degree");
    unsigned L = degree < 0 ? maxdeg : (unsigned)degree;
    d.N = L + 1;
    size_t T = tri(d.N, 0);
    d.C.assign(T, 0.0);
    d.S.assign(T, 0.0);
    bool first = true;
    while (std::getline(f, line)) {
        if (line.empty())
            continue;
        std::istringstream is(line);
        std::string key;
        is >> key;
        if (first) {
            if (key != "gfc")
                throw std::runtime_error("Only static models (gfc) are accepted");
            first = false;
        }
        if (key == "gfc") {
            unsigned l, m;
            double c, s;
            if (is >> l >> m >> c >> s)
                if (l < d.N && m <= l) {
                    d.C[tri(l, m)] = c;
                    d.S[tri(l, m)] = s;
                }
        }
    }
    // gravity_source.cu:76-95 (diagonal and delta) and gravity_source_device.cu:20-46
    d.pdiag.assign(d.N, 0.0);
    d.delta.assign(d.N, 0.0);
    d.pdiag[0] = 1;
    if (d.N > 1) {
        d.pdiag[1] = std::sqrt(3);
        d.delta[1] = std::sqrt(3);
    }
    for (unsigned i = 2; i < d.N; i++) {
        double prev = d.pdiag[i - 1];
        double tmp = 2.0 * (double)i + 1.0;
        d.pdiag[i] = prev * std::sqrt(tmp / (2.0 * (double)i));
        d.delta[i] = prev * std::sqrt(tmp);
    }
    d.xi.assign(T, 0.0);
    d.eta.assign(T, 0.0);
    d.zeta.assign(T, 0.0);
    for (unsigned i = 0; i < d.N; i++)
        for (unsigned j = 0; j < i; j++) {
            double n = i, m = j;
            size_t k = tri(i, j);
            d.xi[k] = std::sqrt((2.0 * n - 1.0) * (2.0 * n + 1.0) / ((n - m) * (n + m)));
            d.eta[k] = std::sqrt((n + m - 1.0) * (2.0 * n + 1.0) * (n - m - 1.0)
                / ((n + m) * (n - m) * (2.0 * n - 3.0)));
            if (m == 0)
                d.zeta[k] = std::sqrt(n * (n + 1.0) / 2.0);
            else
                d.zeta[k] = std::sqrt((n - m) * (n + m + 1.0));
        }
}

// types.cuh:79-130
EopRec eop_get(const std::vector<EopRec>& rec, double t)
{
    size_t size = rec.size();
    if (size == 0)
        return EopRec { -1.0, 0.0, 0.0, 0.0 };
    if (t <= rec[0].mjd)
        return rec[0];
    if (t >= rec[size - 1].mjd)
        return rec[size - 1];
    int left = 0, right = (int)size - 1;
    while (left <= right) {
        int mid = left + (right - left) / 2;
        double cur = rec[mid].mjd;
        if (cur == t)
            return rec[mid];
        else if (cur < t)
            left = mid + 1;
        else
            right = mid - 1;
    }
    const EopRec& r0 = rec[right];
    const EopRec& r1 = rec[left];
    double fr = (t - r0.mjd) / (r1.mjd - r0.mjd);
    EopRec o;
    o.mjd = t;
    o.xp = r0.xp + fr * (r1.xp - r0.xp);
    o.yp = r0.yp + fr * (r1.yp - r0.yp);
    o.dut1 = r0.dut1 + fr * (r1.dut1 - r0.dut1);
    return o;
}

// types.cuh:140-152
double leap_get(const std::vector<double>& mjd, const std::vector<double>& val, double t)
{
    if (mjd.empty())
        return 0.0;
    for (size_t i = mjd.size(); i-- > 0;)
        if (t >= mjd[i])
            return val[i];
    return val[0];
}

// ---------------------------------------------------------------------------------------
// Frames: ref_sys_trans.cu
// ---------------------------------------------------------------------------------------
size_t sq(size_t r, size_t c) { return c * 3 + r; } // gpu_utils.cuh:60 (column-major)

void rot_x(double a, double* p)
{
    double c = std::cos(a), s = std::sin(a);
    p[sq(0, 0)] = 1.0; p[sq(0, 1)] = 0.0; p[sq(0, 2)] = 0.0;
    p[sq(1, 0)] = 0.0; p[sq(1, 1)] = c;   p[sq(1, 2)] = s;
    p[sq(2, 0)] = 0.0; p[sq(2, 1)] = -s;  p[sq(2, 2)] = c;
}
void rot_y(double a, double* p)
{
    double c = std::cos(a), s = std::sin(a);
    p[sq(0, 0)] = c;   p[sq(0, 1)] = 0.0; p[sq(0, 2)] = -s;
    p[sq(1, 0)] = 0.0; p[sq(1, 1)] = 1.0; p[sq(1, 2)] = 0.0;
    p[sq(2, 0)] = s;   p[sq(2, 1)] = 0.0; p[sq(2, 2)] = c;
}
void rot_z(double a, double* p)
{
    double c = std::cos(a), s = std::sin(a);
    p[sq(0, 0)] = c;   p[sq(0, 1)] = s;   p[sq(0, 2)] = 0.0;
    p[sq(1, 0)] = -s;  p[sq(1, 1)] = c;   p[sq(1, 2)] = 0.0;
    p[sq(2, 0)] = 0.0; p[sq(2, 1)] = 0.0; p[sq(2, 2)] = 1.0;
}

double calendar_to_mjd(int year, int mon, int day, int h, int min, double sec)
{
    if (mon <= 2) {
        --year;
        mon += 12;
    }
    int B;
    if ((10000L * year + 100L * mon + day) <= 15821004L)
        B = -2 + ((year + 4716) / 4) - 1179;
    else if ((10000L * year + 100L * mon + day) >= 15821015L)
        B = (year / 400) - (year / 100) + (year / 4);
    else
        return -1.0;
    long mjd = 365L * year - 679004L + B + int(30.6001 * (mon + 1)) + day;
    double frac = (h + min / 60.0 + sec / 3600.0) / 24.0;
    return mjd + frac;
}

void ftable(double t, double* F)
{
    const double l[] = { 134.96340251 * 3600.0, 1717915923.2178, 31.8792, 0.051635, -0.00024470 };
    const double lp[] = { 357.52910918 * 3600.0, 129596581.0481, -0.5532, 0.000136, -0.00001149 };
    const double f[] = { 93.27209062 * 3600.0, 1739527262.8478, -12.7512, -0.001037, 0.00000417 };
    const double dd[] = { 297.85019547 * 3600.0, 1602961601.2090, -6.3706, 0.006593, -0.00003169 };
    const double om[] = { 125.04455501 * 3600.0, -6962890.5431, 7.4722, 0.007702, -0.00005939 };
    F[0] = l[4]; F[1] = lp[4]; F[2] = f[4]; F[3] = dd[4]; F[4] = om[4];
    for (int i = 3; i >= 0; --i) {
        F[0] = F[0] * t + l[i];
        F[1] = F[1] * t + lp[i];
        F[2] = F[2] * t + f[i];
        F[3] = F[3] * t + dd[i];
        F[4] = F[4] * t + om[i];
    }
    for (int i = 0; i < 5; i++)
        F[i] /= kArcsPerRad;
    const double pl[9][3] = { { 4.402608842, 2608.7903141574, 0.0 },
        { 3.176146697, 1021.3285546211, 0.0 }, { 1.753470314, 628.3075849991, 0.0 },
        { 6.203480913, 334.0612426700, 0.0 }, { 0.599546497, 52.9690962641, 0.0 },
        { 0.874016757, 21.3299104960, 0.0 }, { 5.481293872, 7.4781598567, 0.0 },
        { 5.311886287, 3.8133035638, 0.0 }, { 0.0, 0.02438175, 0.00000538691 } };
    for (int q = 0; q < 9; q++)
        F[5 + q] = pl[q][2];
    for (int i = 1; i >= 0; --i)
        for (int q = 0; q < 9; q++)
            F[5 + q] = F[5 + q] * t + pl[q][i];
    for (int i = 0; i < 14; i++)
        F[i] = std::fmod(F[i], 2.0 * M_PI);
}

double nut_delta(const double* F, const std::vector<NutTerm>& terms)
{
    double result = 0;
    for (size_t i = terms.size(); i-- > 0;) {
        double arg = 0;
        for (int k = 13; k >= 0; k--)
            arg += F[k] * terms[i].arg[k];
        result += terms[i].s / (kArcsPerRad * 1e6) * std::sin(arg);
        result += terms[i].c / (kArcsPerRad * 1e6) * std::cos(arg);
    }
    return result;
}

void precession(double t, double& epsA, double& gam, double& phi, double& psi)
{
    const double e[] = { 84381.406, -46.836769, -0.0001831, 0.00200340, -0.000000576,
        -0.0000000434 };
    const double g[] = { -0.052928, 10.556378, 0.4932044, -0.00031238, -0.000002788,
        0.0000000260 };
    const double p[] = { 84381.412819, -46.811016, 0.0511268, 0.00053289, -0.000000440,
        -0.0000000176 };
    const double s[] = { -0.041775, 5038.481484, 1.5584175, -0.00018522, -0.000026452,
        -0.0000000148 };
    epsA = e[5]; gam = g[5]; phi = p[5]; psi = s[5];
    for (int i = 4; i >= 0; --i) {
        epsA = epsA * t + e[i];
        gam = gam * t + g[i];
        phi = phi * t + p[i];
        psi = psi * t + s[i];
    }
    epsA /= kArcsPerRad;
    gam /= kArcsPerRad;
    phi /= kArcsPerRad;
    psi /= kArcsPerRad;
}

double gast(double mjd_ut1, double mjd_tt, double dpsi, double epsA, const double* F)
{
    double T = (mjd_tt - kMjdJ2000) / 36525.0;
    double Du = mjd_ut1 - kMjdJ2000;
    double Dfrac = Du - std::floor(Du);
    double GMST = 86400.0 * (0.7790572732640 + 0.00273781191135448 * Du + Dfrac);
    const double c[6] = { 0.00096707, 307.47710227, 0.092772113, -0.0000000293, -0.00000199707,
        -2.453e-9 };
    double poly = c[5];
    for (int i = 4; i >= 0; --i)
        poly = poly * T + c[i];
    GMST += poly;
    GMST *= (M_PI / 43200.0);
    double l = F[0], lp = F[1], f = F[2], D = F[3], Om = F[4];
    double ee = 0.0;
    ee -= 0.00000087 * T * std::sin(Om);
    ee -= 0.00000063 * std::sin(l - Om);
    ee -= 0.00000063 * std::sin(l + Om);
    ee -= 0.00000126 * std::sin(lp - Om);
    ee -= 0.00000141 * std::sin(lp + Om);
    ee -= 0.00000172 * std::sin(3.0 * Om);
    ee += 0.00000198 * std::sin(2.0 * f + Om);
    ee += 0.00000202 * std::sin(2.0 * f + 3.0 * Om);
    ee -= 0.00000455 * std::sin(2.0 * f - 2.0 * D + 2.0 * Om);
    ee += 0.00001121 * std::sin(2.0 * f - 2.0 * D + Om);
    ee += 0.00001175 * std::sin(2.0 * f - 2.0 * D + 3.0 * Om);
    ee += 0.00006352 * std::sin(2.0 * Om);
    ee += 0.00264096 * std::sin(Om);
    ee *= (M_PI / 43200.0) / 15.0;
    ee += dpsi * std::cos(epsA);
    return std::fmod(GMST + ee, 2 * M_PI);
}

// column-major 3x3 product, summed k = 0,1,2 (time_points.cu:146-152)
void mat3(const double* A, const double* B, double* C)
{
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k)
                s += A[k * 3 + i] * B[j * 3 + k];
            C[j * 3 + i] = s;
        }
}

// TimePointParamsKernel (time_points.cu:26-71) + ComputeITRStoGCRS (:160-201) for one node.
void time_node(const Data& d, const int* ymdhm, double sec, double t, double* M, double* mjd)
{
    double mjd_utc = calendar_to_mjd(ymdhm[0], ymdhm[1], ymdhm[2], ymdhm[3], ymdhm[4], sec + t);
    *mjd = mjd_utc;
    EopRec er = eop_get(d.eop, mjd_utc);
    double mjd_ut1 = mjd_utc + er.dut1 / 86400.0;
    double mjd_tt = mjd_utc + (leap_get(d.ls_mjd, d.ls_val, mjd_utc) + kTtTai) / 86400.0;
    double tq = (mjd_tt - kMjdJ2000) / 36525.0;
    double F[14];
    ftable(tq, F);
    double dpsi = 0, deps = 0;
    dpsi += nut_delta(F, d.lon.j0);
    dpsi += nut_delta(F, d.lon.j1) * tq;
    deps += nut_delta(F, d.obl.j0);
    deps += nut_delta(F, d.obl.j1) * tq;
    double epsA, gam, phi, psib;
    precession(tq, epsA, gam, phi, psib);
    double psi = psib + dpsi;
    double eps = epsA + deps;
    double Q1[9], Q2[9], Q3[9], Q4[9], W1[9], W2[9], W3[9], R[9];
    rot_z(-gam, Q1);
    rot_x(-phi, Q2);
    rot_z(psi, Q3);
    rot_x(eps, Q4);
    double tw = (mjd_tt - kMjdJ2000) / 36525.0;
    double sp = -47e-6 * tw / kArcsPerRad;
    rot_z(-sp, W1);
    rot_y(er.xp, W2);
    rot_x(er.yp, W3);
    rot_z(-gast(mjd_ut1, mjd_tt, dpsi, epsA, F), R);
    double A[9], B[9], Q[9], W[9], QR[9];
    mat3(Q1, Q2, A);
    mat3(Q3, Q4, B);
    mat3(A, B, Q);
    mat3(W1, W2, A);
    mat3(A, W3, W);
    mat3(Q, R, QR);
    mat3(QR, W, M);
}

// ---------------------------------------------------------------------------------------
// Gravity: gravity_source_device.cu:63-239 and gravity_source.cu:168-262
// ---------------------------------------------------------------------------------------
void gravity_itrs(const Data& d, const double* pos, double* g, std::vector<double>& P)
{
    const unsigned N = d.N;
    const size_t T = tri(N, 0);
    P.resize(T);
    double x = pos[0], y = pos[1], z = pos[2];
    double r = std::sqrt(x * x + y * y + z * z);
    double eps = z / r;
    // ComputePbarKernel: diagonal from the table, sub-diagonal eps*delta, column recursion
    for (unsigned i = 0; i < N; ++i)
        P[tri(i, i)] = d.pdiag[i];
    for (unsigned m = 0; m + 1 < N; ++m) {
        unsigned n = m + 1;
        P[tri(n, n - 1)] = eps * d.delta[n];
        while (n != N - 1) {
            ++n;
            size_t k = tri(n, m);
            P[k] = d.xi[k] * eps * P[tri(n - 1, m)] - d.eta[k] * P[tri(n - 2, m)];
        }
    }
    // ComputeCmSmKernel
    std::vector<double> Cm(N), Sm(N);
    Cm[0] = 1;
    Sm[0] = 0;
    if (N > 1) {
        Cm[1] = x / r;
        Sm[1] = y / r;
    }
    for (unsigned i = 2; i < N; i++) {
        Cm[i] = Cm[1] * Cm[i - 1] - Sm[1] * Sm[i - 1];
        Sm[i] = Sm[1] * Cm[i - 1] + Cm[1] * Sm[i - 1];
    }
    // PrepareMSumKernel + reduce over m (sequential), per degree n
    std::vector<double> Gn(N, 0.0), Hn(N, 0.0), Jn(N, 0.0), Kn(N, 0.0);
    for (unsigned n = 0; n < N; ++n) {
        double sg = 0, sh = 0, sj = 0, sk = 0;
        for (unsigned m = 0; m <= n; ++m) {
            size_t k = tri(n, m);
            double G, H, J, K;
            if (n < 2) {
                G = H = J = K = 0;
            } else if (m == 0) {
                double Pn0 = P[tri(n, 0)];
                double Pn1 = P[tri(n, 1)];
                double Cn0 = d.C[tri(n, 0)];
                J = 0;
                K = 0;
                G = Cn0 * (n + 1) * Pn0;
                H = d.zeta[tri(n, 0)] * Cn0 * Pn1;
            } else {
                double Pnm = P[k];
                double Pnm1 = n > m ? P[tri(n, m + 1)] : 0;
                double Cnm = d.C[k], Snm = d.S[k];
                double cm = Cm[m], sm = Sm[m], cm1 = Cm[m - 1], sm1 = Sm[m - 1];
                J = (double)m * Pnm * (Cnm * cm1 + Snm * sm1);
                K = -(double)m * Pnm * (Cnm * sm1 - Snm * cm1);
                G = (double)(1 + n + m) * Pnm * (Cnm * cm + Snm * sm);
                H = d.zeta[k] * Pnm1 * (Cnm * cm + Snm * sm);
            }
            sg += G;
            sh += H;
            sj += J;
            sk += K;
        }
        Gn[n] = sg;
        Hn[n] = sh;
        Jn[n] = sj;
        Kn[n] = sk;
    }
    // PrepareNSumKernel + reduce over n
    double base = d.R / r;
    double coeff = base * base;
    for (unsigned n = 2; n < N; n++) {
        Gn[n] *= coeff;
        Hn[n] *= coeff;
        Kn[n] *= coeff;
        Jn[n] *= coeff;
        coeff *= base;
    }
    double G = 0, H = 0, J = 0, K = 0;
    for (unsigned n = 0; n < N; ++n) {
        G += Gn[n];
        H += Hn[n];
        J += Jn[n];
        K += Kn[n];
    }
    // CalculateAcceleration
    G = G + 1;
    double Lambda = G + z * H / r;
    double ax = Lambda * x / r - J;
    double ay = Lambda * y / r - K;
    double az = Lambda * z / r - H;
    double c = -d.GM / (r * r);
    g[0] = c * ax;
    g[1] = c * ay;
    g[2] = c * az;
}

// a_GCRS = M * g(M^T r): gravity_source.cu:176-178 and :208-210
void gravity_gcrs(const Data& d, const double* M, const double* r, double* a, std::vector<double>& P)
{
    double ri[3], gi[3];
    for (int i = 0; i < 3; ++i) {
        double s = 0.0;
        for (int j = 0; j < 3; ++j)
            s += M[i * 3 + j] * r[j];
        ri[i] = s;
    }
    gravity_itrs(d, ri, gi, P);
    for (int i = 0; i < 3; ++i) {
        double s = 0.0;
        for (int j = 0; j < 3; ++j)
            s += M[j * 3 + i] * gi[j];
        a[i] = s;
    }
}

// ---------------------------------------------------------------------------------------
// Force models: the gravity force model (force_model.cu:47-109) and the analytic test
// forces (unit_tests/include/gpu_test_helpers.cuh, mock_force_model.cuh:69-73, :99)
// ---------------------------------------------------------------------------------------
// ---------------------------------------------------------------------------------------
// Acceleration sources beyond gravity (the extension point of iacceleration_source.cuh:19-33,
// summed by force_model.cu:73-86): Sun and Moon point masses at the low-precision positions of
// Montenbruck and Gill (2000) eq. 3.43-3.49 (J2000 ecliptic, rotated to the J2000 equator), and
// cannonball drag in an exponential atmosphere (Vallado, Table 8-4) rotating with the Earth.
// Added after gravity in the order Sun, Moon, drag.
// ---------------------------------------------------------------------------------------
int g_src_mask = 0;
double g_src_bc = 0.0;
constexpr double kGmSun = 1.32712440018e20, kGmMoon = 4.9028e12, kOmegaE = 7.292115e-5;
constexpr double kAtmR = 6378137.0;
const double kAtm[28][3] = { { 0, 1.225, 7.249 }, { 25, 3.899e-2, 6.349 },
    { 30, 1.774e-2, 6.682 }, { 40, 3.972e-3, 7.554 }, { 50, 1.057e-3, 8.382 },
    { 60, 3.206e-4, 7.714 }, { 70, 8.770e-5, 6.549 }, { 80, 1.905e-5, 5.799 },
    { 90, 3.396e-6, 5.382 }, { 100, 5.297e-7, 5.877 }, { 110, 9.661e-8, 7.263 },
    { 120, 2.438e-8, 9.473 }, { 130, 8.484e-9, 12.636 }, { 140, 3.845e-9, 16.149 },
    { 150, 2.070e-9, 22.523 }, { 180, 5.464e-10, 29.740 }, { 200, 2.789e-10, 37.105 },
    { 250, 7.248e-11, 45.546 }, { 300, 2.418e-11, 53.628 }, { 350, 9.518e-12, 53.298 },
    { 400, 3.725e-12, 58.515 }, { 450, 1.585e-12, 60.828 }, { 500, 6.967e-13, 63.822 },
    { 600, 1.454e-13, 71.835 }, { 700, 3.614e-14, 88.667 }, { 800, 1.170e-14, 124.64 },
    { 900, 5.245e-15, 181.05 }, { 1000, 3.019e-15, 268.00 } };

void ecliptic_to_equator(double lon, double lat, double dist, double* o)
{
    const double eps = 23.43929111 * kPi / 180.0;
    const double xe = dist * std::cos(lat) * std::cos(lon);
    const double ye = dist * std::cos(lat) * std::sin(lon);
    const double ze = dist * std::sin(lat);
    o[0] = xe;
    o[1] = std::cos(eps) * ye - std::sin(eps) * ze;
    o[2] = std::sin(eps) * ye + std::cos(eps) * ze;
}

// T: Julian centuries of TT since J2000. out: Sun xyz, Moon xyz (m).
void sun_moon(double T, double* out)
{
    const double deg = kPi / 180.0, sec = deg / 3600.0;
    const double M = deg * (357.5256 + 35999.049 * T);
    const double lsun = deg * 282.94 + M + sec * (6892.0 * std::sin(M) + 72.0 * std::sin(2.0 * M));
    const double rsun = 1e9 * (149.619 - 2.499 * std::cos(M) - 0.021 * std::cos(2.0 * M));
    ecliptic_to_equator(lsun, 0.0, rsun, out);
    const double L0 = deg * (218.31617 + 481267.88088 * T - 1.3972 * T);
    const double l = deg * (134.96292 + 477198.86753 * T);
    const double lp = deg * (357.52543 + 35999.04944 * T);
    const double F = deg * (93.27283 + 483202.01873 * T);
    const double D = deg * (297.85027 + 445267.11135 * T);
    struct Term {
        double amp, cl, clp, cF, cD;
    };
    // longitude perturbations (arcsec) and their arguments l, l', F, D
    const Term lon[14] = { { 22640, 1, 0, 0, 0 }, { 769, 2, 0, 0, 0 }, { -4586, 1, 0, 0, -2 },
        { 2370, 0, 0, 0, 2 }, { -668, 0, 1, 0, 0 }, { -412, 0, 0, 2, 0 }, { -212, 2, 0, 0, -2 },
        { -206, 1, 1, 0, -2 }, { 192, 1, 0, 0, 2 }, { -165, 0, 1, 0, -2 }, { 148, 1, -1, 0, 0 },
        { -125, 0, 0, 0, 1 }, { -110, 1, 1, 0, 0 }, { -55, 0, 0, 2, -2 } };
    double dl = 0.0;
    for (const Term& t : lon)
        dl += t.amp * std::sin(t.cl * l + t.clp * lp + t.cF * F + t.cD * D);
    const double lmoon = L0 + sec * dl;
    const Term lat[7] = { { -526, 0, 0, 1, -2 }, { 44, 1, 0, 1, -2 }, { -31, -1, 0, 1, -2 },
        { -25, -2, 0, 1, 0 }, { -23, 0, 1, 1, -2 }, { 21, -1, 0, 1, 0 }, { 11, 0, -1, 1, -2 } };
    double db = 18520.0 * std::sin(F + lmoon - L0 + sec * (412.0 * std::sin(2.0 * F) + 541.0 * std::sin(lp)));
    for (const Term& t : lat)
        db += t.amp * std::sin(t.cl * l + t.clp * lp + t.cF * F + t.cD * D);
    const double bmoon = sec * db;
    const Term dist[8] = { { -20905, 1, 0, 0, 0 }, { -3699, -1, 0, 0, 2 }, { -2956, 0, 0, 0, 2 },
        { -570, 2, 0, 0, 0 }, { 246, 2, 0, 0, -2 }, { -205, 0, 1, 0, -2 }, { -171, 1, 0, 0, 2 },
        { -152, 1, 1, 0, -2 } };
    double rk = 385000.0;
    for (const Term& t : dist)
        rk += t.amp * std::cos(t.cl * l + t.clp * lp + t.cF * F + t.cD * D);
    ecliptic_to_equator(lmoon, bmoon, 1e3 * rk, out + 3);
}

double density(double h)
{
    const double km = h / 1000.0;
    int b = 0;
    for (int q = 1; q < 28; ++q)
        if (km >= kAtm[q][0])
            b = q;
    return kAtm[b][1] * std::exp((kAtm[b][0] - km) / kAtm[b][2]);
}

void point_mass(double gm, const double* s, const double* r, double* a)
{
    double d[3], dn = 0.0, sn = 0.0;
    for (int i = 0; i < 3; ++i) {
        d[i] = s[i] - r[i];
        dn += d[i] * d[i];
        sn += s[i] * s[i];
    }
    const double d3 = dn * std::sqrt(dn), s3 = sn * std::sqrt(sn);
    for (int i = 0; i < 3; ++i)
        a[i] += gm * (d[i] / d3 - s[i] / s3);
}

void add_sources(int mask, double bc, const double* tb, const double* r, const double* v, double* a)
{
    if (mask & 1)
        point_mass(kGmSun, tb, r, a);
    if (mask & 2)
        point_mass(kGmMoon, tb + 3, r, a);
    if (mask & 4) {
        const double vr[3] = { v[0] + kOmegaE * r[1], v[1] - kOmegaE * r[0], v[2] };
        const double speed = std::sqrt(vr[0] * vr[0] + vr[1] * vr[1] + vr[2] * vr[2]);
        const double rn = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        const double k = -0.5 * bc * density(rn - kAtmR) * speed;
        for (int i = 0; i < 3; ++i)
            a[i] += k * vr[i];
    }
}

struct Force {
    int kind = 1;
    double h = 1;
    const Data* d = nullptr;
    std::shared_ptr<std::vector<double>> genM, propM; // 9 per node, shared between objects
    std::shared_ptr<std::vector<double>> genTB, propTB; // Sun and Moon per node (6)
    int src_mask = 0;
    double src_bc = 0.0;
    std::vector<double> P;                            // per-thread scratch

    static void analytic(int kind, const double* p, const double* v, double t, double* a)
    {
        switch (kind) {
        case 1:
            a[0] = a[1] = a[2] = 0;
            return;
        case 2:
            a[0] = 1;
            a[1] = 2;
            a[2] = 3;
            return;
        case 3: {
            double val = -3 * 2 * t + 7 * 6 * t * t * t * t * t;
            a[0] = a[1] = a[2] = val;
            return;
        }
        case 4: {
            double val = 56 * std::pow(t, 6) - 42 * std::pow(t, 5) + 60 * std::pow(t, 4)
                - 100 * std::pow(t, 3) - 120 * std::pow(t, 2) + 6 * t - 2;
            a[0] = a[1] = a[2] = val;
            return;
        }
        case 5: {
            double val = 72 * std::pow(t, 7) + 56 * std::pow(t, 6) - 42 * std::pow(t, 5)
                + 60 * std::pow(t, 4) - 100 * std::pow(t, 3) - 120 * std::pow(t, 2) + 6 * t - 2;
            a[0] = a[1] = a[2] = val;
            return;
        }
        case 6: {
            double val = std::exp(-t);
            a[0] = a[1] = a[2] = val;
            return;
        }
        case 7:
            for (int i = 0; i < 3; ++i)
                a[i] = -2 * p[i] - 2 * v[i];
            return;
        }
        a[0] = a[1] = a[2] = 0;
    }

    // ForceModel::ComputeGenerationAcceleration
    void gen(const double* x, const double* v, int stepNum, int k, double* a)
    {
        if (kind == 0) {
            int s = stepNum > 0 ? stepNum - 1 : stepNum;
            int idx = (s + 4) * 10 + k;
            gravity_gcrs(*d, &(*genM)[idx * 9], x, a, P);
            if (src_mask)
                add_sources(src_mask, src_bc, &(*genTB)[idx * 6], x, v, a);
            return;
        }
        int dir = stepNum > 0 ? -1 : 1;
        double t = static_cast<double>(stepNum + dir) * h + (-dir) * kRk8Mid[k] * h;
        analytic(kind, x, v, t, a);
    }
    // ForceModel::ComputePropagationAcceleration, node index stepNum (time (stepNum-4)h)
    void prop(const double* x, const double* v, int node, double* a)
    {
        if (kind == 0) {
            gravity_gcrs(*d, &(*propM)[(size_t)node * 9], x, a, P);
            if (src_mask)
                add_sources(src_mask, src_bc, &(*propTB)[(size_t)node * 6], x, v, a);
            return;
        }
        double t = (double)(node - 4) * h;
        analytic(kind, x, v, t, a);
    }
};

// Node offsets (seconds from the epoch) of the 80 RK8 stage nodes, time_points.cu:35-46 with
// beforeEpoch = 4. With quirk bit 1 the backward stages use the reference's mirrored times
// (s + c_k) h; without it they use the true stage times (s + 1 - c_k) h.
std::vector<double> gen_offsets(double h, int quirks)
{
    std::vector<double> t(80);
    for (int idx = 0; idx < 80; ++idx) {
        int blk = idx / 10, k = idx % 10;
        int s = blk - 4; // s' = stepNum-1 (forward) or stepNum (backward)
        if ((quirks & 2) || s >= 0)
            t[idx] = ((double)blk + kRk8Mid[k]) * h;
        else
            t[idx] = ((double)(s + 1) - kRk8Mid[k]) * h + 4 * h;
    }
    return t;
}

void build_force(Force& f, int kind, int quirks, const int* ymdhm, double sec, double h,
    int maxSteps)
{
    f.kind = kind;
    f.h = h;
    if (kind != 0)
        return;
    if (!g_data)
        throw std::runtime_error("datasets not loaded");
    f.d = g_data.get();
    double shifted = sec - (double)4 * h; // Sec - beforeEpoch * stepSize
    std::vector<double> go = gen_offsets(h, quirks);
    f.genM = std::make_shared<std::vector<double>>(80 * 9);
    f.genTB = std::make_shared<std::vector<double>>(80 * 6);
    f.src_mask = g_src_mask;
    f.src_bc = g_src_bc;
    // Sun and Moon at the node's TT (the time-node kernel's T)
    auto tb = [&](double mjd_utc, double* out) {
        const double mjd_tt = mjd_utc + (leap_get(f.d->ls_mjd, f.d->ls_val, mjd_utc) + kTtTai) / 86400.0;
        sun_moon((mjd_tt - kMjdJ2000) / 36525.0, out);
    };
    double mjd;
    for (int i = 0; i < 80; ++i) {
        time_node(*f.d, ymdhm, shifted, go[i], &(*f.genM)[i * 9], &mjd);
        tb(mjd, &(*f.genTB)[i * 6]);
    }
    int len = 4 + maxSteps + 1;
    f.propM = std::make_shared<std::vector<double>>((size_t)len * 9);
    f.propTB = std::make_shared<std::vector<double>>((size_t)len * 6);
    for (int i = 0; i < len; ++i) {
        time_node(*f.d, ymdhm, shifted, i * h, &(*f.propM)[(size_t)i * 9], &mjd);
        tb(mjd, &(*f.propTB)[(size_t)i * 6]);
    }
}

// ---------------------------------------------------------------------------------------
// Startup: gpu_startup.cu:55-312 and src/device/*.cu. Matrices are [9][3] per object.
// ---------------------------------------------------------------------------------------
struct Startup {
    double pos[9][3], vel[9][3], acc[9][3], s[9][3], S[9][3];
};

void rk8_step(Force& f, const double a[9][9], double h, Startup& st, int step)
{
    double hs = step > 0 ? h : -h;
    int prev = 4 + step + (step > 0 ? -1 : 1);
    double Kx[10][3], Kv[10][3], tX[3], tV[3];
    for (int i = 0; i < 3; ++i) {
        tX[i] = st.pos[prev][i];
        tV[i] = st.vel[prev][i];
        Kx[0][i] = tV[i];
    }
    f.gen(tX, tV, step, 0, Kv[0]);
    for (int k = 1; k <= 9; k++) {
        for (int i = 0; i < 3; ++i) {
            double sx = 0.0, sv = 0.0;
            for (int j = 0; j < k; ++j) {
                sx += Kx[j][i] * a[k - 1][j];
                sv += Kv[j][i] * a[k - 1][j];
            }
            tX[i] = hs * sx + tX[i];
            tV[i] = hs * sv + tV[i];
        }
        f.gen(tX, tV, step, k, Kv[k]);
        for (int i = 0; i < 3; ++i) {
            Kx[k][i] = tV[i];
            tX[i] = st.pos[prev][i];
            tV[i] = st.vel[prev][i];
        }
    }
    for (int i = 0; i < 3; ++i) {
        double sv = 0.0, sx = 0.0;
        for (int k = 0; k < 10; ++k) {
            sv += Kv[k][i] * kRk8b[k];
            sx += Kx[k][i] * kRk8b[k];
        }
        st.vel[4 + step][i] = hs * sv + st.vel[prev][i];
        st.pos[4 + step][i] = hs * sx + st.pos[prev][i];
    }
}

void rk8_predict(Force& f, int quirks, double h, const double* r0, const double* v0, Startup& st)
{
    double a[9][9];
    rk8a(quirks, a);
    std::memset(&st, 0, sizeof(st));
    for (int i = 0; i < 3; ++i) {
        st.pos[4][i] = r0[i];
        st.vel[4][i] = v0[i];
    }
    for (int i = 1; i <= 4; i++) {
        rk8_step(f, a, h, st, i);
        rk8_step(f, a, h, st, -i);
    }
}

void acc_matrix(Force& f, Startup& st, bool epoch)
{
    for (int i = 0; i < 9; ++i) {
        if (i == 4 && !epoch)
            continue;
        f.prop(st.pos[i], st.vel[i], i, st.acc[i]);
    }
}

void calc_s0(Startup& st, double h)
{
    double invh = 1 / h;
    double invh2 = 1 / (h * h);
    for (int i = 0; i < 3; ++i) {
        double ss = 0.0, sS = 0.0;
        for (int k = 0; k < 9; ++k) {
            ss += st.acc[k][i] * kSA[4][k];
            sS += st.acc[k][i] * kGJ[4][k];
        }
        st.s[4][i] = invh * st.vel[4][i] + (-ss);
        st.S[4][i] = invh2 * st.pos[4][i] + (-sS);
    }
}

void calc_all_s(Startup& st)
{
    for (int i = 1; i <= 4; i++)
        for (int j = 0; j < 3; j++) {
            st.s[4 + i][j] = st.s[4 + i - 1][j] + (st.acc[4 + i - 1][j] + st.acc[4 + i][j]) / 2;
            st.s[4 - i][j] = st.s[4 - i + 1][j] - (st.acc[4 - i + 1][j] + st.acc[4 - i][j]) / 2;
            st.S[4 + i][j] = st.s[4 + i - 1][j] + st.S[4 + i - 1][j] + st.acc[4 + i - 1][j] / 2;
            st.S[4 - i][j] = -st.s[4 - i + 1][j] + st.S[4 - i + 1][j] + st.acc[4 - i + 1][j] / 2;
        }
}

void correct(Force& f, Startup& st, double h)
{
    calc_s0(st, h);
    calc_all_s(st);
    double ep[3], ev[3];
    for (int i = 0; i < 3; ++i) {
        ep[i] = st.pos[4][i];
        ev[i] = st.vel[4][i];
    }
    double hh = h * h;
    for (int j = 0; j < 9; ++j)
        for (int i = 0; i < 3; ++i) {
            double sv = 0.0, sp = 0.0;
            for (int k = 0; k < 9; ++k) {
                sv += st.acc[k][i] * kSA[j][k];
                sp += st.acc[k][i] * kGJ[j][k];
            }
            st.vel[j][i] = (st.s[j][i] + sv) * h;
            st.pos[j][i] = (st.S[j][i] + sp) * hh;
        }
    for (int i = 0; i < 3; ++i) {
        st.pos[4][i] = ep[i];
        st.vel[4][i] = ev[i];
    }
    acc_matrix(f, st, false);
}

void startup(Force& f, int quirks, double h, int ncorr, const double* r0, const double* v0,
    Startup& st)
{
    rk8_predict(f, quirks, h, r0, v0, st);
    acc_matrix(f, st, true);
    for (int i = 0; i < ncorr; i++)
        correct(f, st, h);
}

// ---------------------------------------------------------------------------------------
// GJ8 step: GpuPropagator::Step / Correct (gpu_propagator.cu:28-119), Predictor::Predict
// (prediction.cu:49-74)
// ---------------------------------------------------------------------------------------
struct GJState {
    double r[3], v[3], s[3], S[3], acc[9][3];
};

void gj_predict_core(GJState& g, double h)
{
    for (int i = 0; i < 3; ++i) {
        double sv = 0.0, sp = 0.0;
        for (int k = 0; k < 9; ++k) {
            sv += g.acc[k][i] * kSA[9][k];
            sp += g.acc[k][i] * kGJ[9][k];
        }
        g.S[i] += g.s[i] + g.acc[8][i] / 2;
        g.v[i] = (sv + g.s[i] + g.acc[8][i] / 2) * h;
        g.r[i] = (sp + g.S[i]) * h * h;
    }
}

void gj_finalize(double* r, double* v, const double* a8, const double* s, const double* S, double h)
{
    for (int i = 0; i < 3; ++i) {
        v[i] = h * (v[i] + a8[i] * kSA[8][8] + s[i]);
        r[i] = h * h * (r[i] + a8[i] * kGJ[8][8] + S[i]);
    }
}

void gj_step(Force& f, GJState& g, double h, int stepNum, int ncorr)
{
    gj_predict_core(g, h);
    for (int n = 0; n < 8; n++)
        for (int i = 0; i < 3; ++i)
            g.acc[n][i] = g.acc[n + 1][i];
    f.prop(g.r, g.v, stepNum + 4, g.acc[8]);
    double sprev[3], T1[3], T2[3];
    for (int i = 0; i < 3; ++i) {
        sprev[i] = g.s[i];
        double a = 0.0, b = 0.0;
        for (int k = 0; k < 8; ++k) {
            a += g.acc[k][i] * kSA[8][k];
            b += g.acc[k][i] * kGJ[8][k];
        }
        T1[i] = a;
        T2[i] = b;
    }
    for (int it = 0; it < ncorr; it++) {
        // The first corrector evaluation sees exactly the predicted state, so it would
        // reproduce the predictor's evaluation bit for bit; it is reused instead of redone.
        if (it > 0)
            f.prop(g.r, g.v, stepNum + 4, g.acc[8]);
        for (int i = 0; i < 3; ++i) {
            g.s[i] = sprev[i] + (g.acc[8][i] + g.acc[7][i]) / 2;
            g.v[i] = T1[i];
            g.r[i] = T2[i];
        }
        gj_finalize(g.r, g.v, g.acc[8], g.s, g.S, h);
    }
}

// One object through `steps` GpuPropagator::Step calls.
void propagate_one(Force& f, int quirks, double h, int steps, int ncorr, const double* r0,
    const double* v0, double* rout, double* vout, double* traj, long nobj, long obj)
{
    Startup st;
    GJState g;
    if (steps <= 0) {
        for (int i = 0; i < 3; ++i) {
            rout[i] = r0[i];
            vout[i] = v0[i];
        }
        return;
    }
    startup(f, quirks, h, ncorr, r0, v0, st);
    for (int i = 0; i < 3; ++i) {
        g.s[i] = st.s[8][i];
        g.S[i] = st.S[8][i];
    }
    std::memcpy(g.acc, st.acc, sizeof(g.acc));
    for (int step = 1; step <= steps; ++step) {
        if (step <= 4) {
            for (int i = 0; i < 3; ++i) {
                g.r[i] = st.pos[step + 4][i];
                g.v[i] = st.vel[step + 4][i];
            }
        } else {
            gj_step(f, g, h, step, ncorr);
        }
        if (traj) {
            double* o = traj + ((size_t)(step - 1) * nobj + obj) * 9;
            for (int i = 0; i < 3; ++i) {
                o[i] = g.r[i];
                o[3 + i] = g.v[i];
                o[6 + i] = g.acc[8][i];
            }
        }
    }
    for (int i = 0; i < 3; ++i) {
        rout[i] = g.r[i];
        vout[i] = g.v[i];
    }
}

// ---------------------------------------------------------------------------------------
// Conjunction: partition_parameters.cuh:23-143, interpolation.cuh:27-152, device.cu:56-173
// ---------------------------------------------------------------------------------------
struct Partition {
    std::vector<unsigned> ids, offsets, higher, mapoff;
    unsigned numSets = 0, numPairs = 0;
    bool calc(const unsigned* set, int n)
    {
        ids.assign(set, set + n);
        offsets = { 0 };
        mapoff = { 0 };
        std::vector<unsigned> sizes;
        higher.clear();
        unsigned prevId = set[0], prevChange = 0;
        for (unsigned i = 1; i < (unsigned)n; i++) {
            if (set[i] < prevId || set[i] > prevId + 1)
                return false;
            if (set[i] > prevId) {
                offsets.push_back(i);
                sizes.push_back(i - prevChange);
                higher.push_back((unsigned)n - i);
                mapoff.push_back(mapoff.back() + sizes.back() * higher.back());
                prevChange = i;
            }
            prevId = set[i];
        }
        sizes.push_back((unsigned)n - prevChange);
        higher.push_back(0);
        numSets = (unsigned)offsets.size();
        numPairs = mapoff.back();
        return true;
    }
    int idx(unsigned a, unsigned b) const
    {
        if (a > b)
            std::swap(a, b);
        unsigned setA = ids[a];
        if (setA == numSets - 1)
            return -1;
        int o1 = (int)b - (int)offsets[setA + 1];
        if (o1 < 0)
            return -1;
        return o1 + (int)mapoff[setA] + (int)((a - offsets[setA]) * higher[setA]);
    }
    bool pair(unsigned i, unsigned& a, unsigned& b) const
    {
        unsigned setA = 0;
        while (setA < numSets - 1 && mapoff[setA + 1] <= i)
            setA++;
        if (setA >= numSets - 1)
            return false;
        a = (i - mapoff[setA]) / higher[setA] + offsets[setA];
        b = (i - mapoff[setA]) % higher[setA] + offsets[setA + 1];
        return true;
    }
};

struct Hermite {
    double c[9], dt;
    Hermite(const double* q, double dt_)
        : dt(dt_)
    {
        double p0 = q[0], v0 = q[1], a0 = q[2], p1 = q[3], v1 = q[4], a1 = q[5], p2 = q[6],
               v2 = q[7], a2 = q[8];
        double t[9] = { 0.0, 0.0, 0.0, dt, dt, dt, 2.0 * dt, 2.0 * dt, 2.0 * dt };
        c[0] = c[1] = c[2] = p0;
        c[3] = c[4] = c[5] = p1;
        c[6] = c[7] = c[8] = p2;
        for (int j = 1; j < 9; j++)
            for (int i = 8; i >= j; i--) {
                if (t[i] == t[i - j]) {
                    if (j == 1)
                        c[i] = (i < 3) ? v0 : (i < 6) ? v1 : v2;
                    else if (j == 2)
                        c[i] = ((i < 3) ? a0 : (i < 6) ? a1 : a2) / 2.0;
                } else {
                    c[i] = (c[i] - c[i - 1]) / (t[i] - t[i - j]);
                }
            }
    }
    double val(double t) const
    {
        double z[9] = { 0.0, 0.0, 0.0, dt, dt, dt, 2.0 * dt, 2.0 * dt, 2.0 * dt };
        double res = c[0], term = 1.0;
        for (int i = 1; i < 9; ++i) {
            term *= (t - z[i - 1]);
            res += c[i] * term;
        }
        return res;
    }
    double der(double te) const
    {
        double z[9] = { 0.0, 0.0, 0.0, dt, dt, dt, 2.0 * dt, 2.0 * dt, 2.0 * dt };
        double res = 0.0, Nn = 1.0, dN = 0.0;
        for (int i = 0; i < 9; ++i) {
            res += c[i] * dN;
            if (i < 8) {
                dN = dN * (te - z[i]) + Nn;
                Nn *= (te - z[i]);
            }
        }
        return res;
    }
};

// ApproachFunc (interpolation.cuh:100-125): f(t) = (A - B)(t) . (A - B)'(t) for two objects
// interpolated per axis.
struct Approach {
    const Hermite *A[3], *B[3];
    double operator()(double t) const
    {
        double dx = A[0]->val(t) - B[0]->val(t), dy = A[1]->val(t) - B[1]->val(t),
               dz = A[2]->val(t) - B[2]->val(t);
        double ux = A[0]->der(t) - B[0]->der(t), uy = A[1]->der(t) - B[1]->der(t),
               uz = A[2]->der(t) - B[2]->der(t);
        return (dx * ux) + (dy * uy) + (dz * uz);
    }
};

// FindRoot (interpolation.cuh:127-152): secant iteration from (x0, x1); a flat secant returns x1.
template <class F> double find_root(F f, double x0, double x1, double tol = 1e-7, int maxit = 100)
{
    double f0 = f(x0), f1 = f(x1);
    for (int i = 0; i < maxit; ++i) {
        if (std::fabs(f1 - f0) < 1e-14)
            return x1;
        double xn = x1 - f1 * (x1 - x0) / (f1 - f0);
        double fn = f(xn);
        if (std::fabs(xn - x1) < tol || std::fabs(fn) < tol)
            return xn;
        x0 = x1;
        f0 = f1;
        x1 = xn;
        f1 = fn;
    }
    return x1;
}

// Squared distance with an explicit fused order; the GPU screen uses the identical sequence,
// so identical trajectories give identical bits.
inline double dist_of(const double* a, const double* b)
{
    double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
    return std::sqrt(std::fma(dz, dz, std::fma(dy, dy, dx * dx)));
}

} // namespace

// =========================================================================================
// C ABI
// =========================================================================================
extern "C" {

int orc_open(const char* dir, int degree)
{
    try {
        auto d = std::make_unique<Data>();
        std::string p(dir);
        load_gravity(p + "/JGM3.gfc", degree, *d);
        d->lon = load_nutation(p + "/IAU2000_R06_NutLon.txt");
        d->obl = load_nutation(p + "/IAU2000_R06_NutObl.txt");
        d->eop = load_eop(p + "/IAU2000A_EOP.txt");
        load_leap(p + "/leap-seconds.list", d->ls_mjd, d->ls_val);
        if (d->lon.j0.empty() || d->obl.j0.empty())
            throw std::runtime_error("nutation tables are empty");
        g_data = std::move(d);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}
void orc_close(void) { g_data.reset(); }
const char* orc_error(void) { return g_err.c_str(); }
int orc_degree(void) { return g_data ? (int)g_data->N - 1 : -1; }
double orc_gm(void) { return g_data ? g_data->GM : 0; }
double orc_radius(void) { return g_data ? g_data->R : 0; }

double orc_calendar_to_mjd(int y, int mo, int d, int h, int mi, double sec)
{
    return calendar_to_mjd(y, mo, d, h, mi, sec);
}
double orc_ntp_to_mjd(long long ntp) { return kMjd1900 + (static_cast<double>(ntp) / 86400.0); }
void orc_ftable(double t, double* F) { ftable(t, F); }
void orc_precession(double t, double* o) { precession(t, o[0], o[1], o[2], o[3]); }
void orc_nutation(double t, double* o)
{
    double F[14];
    ftable(t, F);
    o[0] = 0;
    o[1] = 0;
    if (!g_data)
        return;
    o[0] += nut_delta(F, g_data->lon.j0);
    o[0] += nut_delta(F, g_data->lon.j1) * t;
    o[1] += nut_delta(F, g_data->obl.j0);
    o[1] += nut_delta(F, g_data->obl.j1) * t;
}
double orc_gast(double u, double t, double dpsi, double eps, const double* F)
{
    return gast(u, t, dpsi, eps, F);
}
void orc_rot(int axis, double a, double* M)
{
    if (axis == 0)
        rot_x(a, M);
    else if (axis == 1)
        rot_y(a, M);
    else
        rot_z(a, M);
}
void orc_eop(double mjd, double* o)
{
    EopRec r = g_data ? eop_get(g_data->eop, mjd) : EopRec { -1, 0, 0, 0 };
    o[0] = r.mjd;
    o[1] = r.xp;
    o[2] = r.yp;
    o[3] = r.dut1;
}
double orc_leap_seconds(double mjd)
{
    return g_data ? leap_get(g_data->ls_mjd, g_data->ls_val, mjd) : 0.0;
}
int orc_time_nodes(const int* ymdhm, double sec, const double* times, int len, double* M, double* mjd)
{
    if (!g_data) {
        g_err = "datasets not loaded";
        return 1;
    }
    for (int i = 0; i < len; ++i)
        time_node(*g_data, ymdhm, sec, times[i], M + 9 * i, mjd + i);
    return 0;
}

int orc_gravity_itrs(const double* r, long n, double* a)
{
    if (!g_data) {
        g_err = "datasets not loaded";
        return 1;
    }
    std::vector<double> P;
    for (long i = 0; i < n; ++i)
        gravity_itrs(*g_data, r + 3 * i, a + 3 * i, P);
    return 0;
}

double orc_potential_itrs(const double* pos)
{
    // Textbook spherical-harmonic potential in long double:
    // U = GM/r sum_n (R/r)^n sum_m Pbar_nm(sin phi) (C cos m lam + S sin m lam), with the
    // fully normalised associated Legendre functions by the standard column recursion.
    if (!g_data)
        return 0.0;
    const Data& d = *g_data;
    typedef long double ld;
    ld x = pos[0], y = pos[1], z = pos[2];
    ld r = sqrtl(x * x + y * y + z * z);
    ld sphi = z / r, cphi = sqrtl(x * x + y * y) / r;
    ld lam = atan2l(y, x);
    unsigned N = d.N;
    std::vector<ld> P((size_t)N * N, 0.0L);
    auto at = [&](unsigned n, unsigned m) -> ld& { return P[(size_t)n * N + m]; };
    at(0, 0) = 1.0L;
    if (N > 1)
        at(1, 1) = sqrtl(3.0L) * cphi;
    for (unsigned m = 2; m < N; ++m)
        at(m, m) = sqrtl((2.0L * m + 1.0L) / (2.0L * m)) * cphi * at(m - 1, m - 1);
    for (unsigned m = 0; m < N; ++m) {
        if (m + 1 < N)
            at(m + 1, m) = sqrtl(2.0L * m + 3.0L) * sphi * at(m, m);
        for (unsigned n = m + 2; n < N; ++n) {
            ld a = sqrtl((2.0L * n - 1.0L) * (2.0L * n + 1.0L) / ((ld)(n - m) * (ld)(n + m)));
            ld b = sqrtl((2.0L * n + 1.0L) * (ld)(n + m - 1) * (ld)(n - m - 1)
                / ((ld)(n - m) * (ld)(n + m) * (2.0L * n - 3.0L)));
            at(n, m) = a * sphi * at(n - 1, m) - b * at(n - 2, m);
        }
    }
    ld sum = 0.0L;
    ld rr = (ld)d.R / r, rn = rr;
    for (unsigned n = 2; n < N; ++n) {
        rn = powl(rr, (ld)n);
        ld sn = 0.0L;
        for (unsigned m = 0; m <= n; ++m)
            sn += at(n, m) * ((ld)d.C[tri(n, m)] * cosl(m * lam) + (ld)d.S[tri(n, m)] * sinl(m * lam));
        sum += rn * sn;
    }
    return (double)((ld)d.GM / r * sum); // non-central part only
}

void orc_gj_predict(const double* acc27, const double* s3, double* S3, double h, double* r3, double* v3)
{
    GJState g;
    std::memcpy(g.acc, acc27, sizeof(g.acc));
    for (int i = 0; i < 3; ++i) {
        g.s[i] = s3[i];
        g.S[i] = S3[i];
    }
    gj_predict_core(g, h);
    for (int i = 0; i < 3; ++i) {
        S3[i] = g.S[i];
        r3[i] = g.r[i];
        v3[i] = g.v[i];
    }
}

void orc_gj_finalize_correction(
    double* r3, double* v3, const double* acc27, const double* s3, const double* S3, double h)
{
    gj_finalize(r3, v3, acc27 + 24, s3, S3, h);
}

int orc_startup_piece(int mode, int kind, double h, double* pos27, double* vel27, double* acc27,
    double* s27, double* S27)
{
    try {
        Force f;
        int ymdhm[5] = { 2000, 1, 1, 0, 0 };
        build_force(f, kind, 3, ymdhm, 0.0, h, 4);
        Startup st;
        std::memset(&st, 0, sizeof(st));
        std::memcpy(st.pos, pos27, sizeof(st.pos));
        std::memcpy(st.vel, vel27, sizeof(st.vel));
        acc_matrix(f, st, true);
        if (mode == 0)
            calc_s0(st, h);
        else if (mode == 1)
            correct(f, st, h);
        std::memcpy(pos27, st.pos, sizeof(st.pos));
        std::memcpy(vel27, st.vel, sizeof(st.vel));
        std::memcpy(acc27, st.acc, sizeof(st.acc));
        std::memcpy(s27, st.s, sizeof(st.s));
        std::memcpy(S27, st.S, sizeof(st.S));
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

int orc_rk8_nodes(int kind, int quirks, const int* ymdhm, double sec, double h, const double* r0,
    const double* v0, double* pos27, double* vel27)
{
    try {
        Force f;
        build_force(f, kind, quirks, ymdhm, sec, h, 4);
        Startup st;
        rk8_predict(f, quirks, h, r0, v0, st);
        std::memcpy(pos27, st.pos, sizeof(st.pos));
        std::memcpy(vel27, st.vel, sizeof(st.vel));
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

int orc_startup(int kind, int quirks, const int* ymdhm, double sec, double h, int ncorr,
    const double* r0, const double* v0, double* pos27, double* vel27, double* s27, double* S27,
    double* acc27)
{
    try {
        Force f;
        build_force(f, kind, quirks, ymdhm, sec, h, 4);
        Startup st;
        startup(f, quirks, h, ncorr, r0, v0, st);
        std::memcpy(pos27, st.pos, sizeof(st.pos));
        std::memcpy(vel27, st.vel, sizeof(st.vel));
        std::memcpy(s27, st.s, sizeof(st.s));
        std::memcpy(S27, st.S, sizeof(st.S));
        std::memcpy(acc27, st.acc, sizeof(st.acc));
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

int orc_propagate(int kind, int quirks, const int* ymdhm, double sec, long n, const double* r0,
    const double* v0, const double* h, const int* steps, int ncorr, double* rout, double* vout,
    double* traj, int threads)
{
    try {
        if (n <= 0)
            return 0;
        // one force model (time-node tables) per distinct step size
        std::map<double, int> maxSteps;
        for (long i = 0; i < n; ++i) {
            if (!(h[i] > 0))
                throw std::runtime_error("step size must be positive");
            int& m = maxSteps[h[i]];
            m = std::max(m, steps[i]);
        }
        std::map<double, std::shared_ptr<Force>> forces;
        for (auto& kv : maxSteps) {
            auto f = std::make_shared<Force>();
            build_force(*f, kind, quirks, ymdhm, sec, kv.first, kv.second);
            forces[kv.first] = f;
        }
#ifdef _OPENMP
        int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
#endif
        for (long i = 0; i < n; ++i) {
            Force f = *forces.find(h[i])->second; // private scratch, node tables shared
            propagate_one(f, quirks, h[i], steps[i], ncorr, r0 + 3 * i, v0 + 3 * i,
                rout + 3 * i, vout + 3 * i, traj, n, i);
        }
        (void)threads;
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

int orc_partition(const unsigned* set, int n, unsigned* so, unsigned* nh, unsigned* mo,
    unsigned* np)
{
    Partition p;
    if (n <= 0 || !p.calc(set, n))
        return -1;
    for (unsigned i = 0; i < p.numSets; ++i) {
        so[i] = p.offsets[i];
        nh[i] = p.higher[i];
        mo[i] = p.mapoff[i];
    }
    *np = p.numPairs;
    return (int)p.numSets;
}
int orc_partition_idx(const unsigned* set, int n, unsigned a, unsigned b)
{
    Partition p;
    if (!p.calc(set, n))
        return -2;
    return p.idx(a, b);
}
int orc_partition_pair(const unsigned* set, int n, unsigned i, unsigned* a, unsigned* b)
{
    Partition p;
    if (!p.calc(set, n))
        return -2;
    return p.pair(i, *a, *b) ? 1 : 0;
}

void orc_hermite(const double* q, double dt, double* c)
{
    Hermite hp(q, dt);
    std::memcpy(c, hp.c, sizeof(hp.c));
}
double orc_hermite_eval(const double* c, double dt, double t, int deriv)
{
    double q[9] = { 0 };
    Hermite hp(q, dt);
    std::memcpy(hp.c, c, sizeof(hp.c));
    return deriv ? hp.der(t) : hp.val(t);
}

double orc_find_root_poly(const double* coef, int ncoef, double x0, double x1, double tol)
{
    auto f = [&](double x) {
        double y = 0.0;
        for (int i = ncoef - 1; i >= 0; --i)
            y = y * x + coef[i];
        return y;
    };
    return find_root(f, x0, x1, tol);
}

double orc_approach_func(const double* qa, const double* qb, double dt, double t, double x1,
    double tol, int root)
{
    const Hermite A[3] = { Hermite(qa, dt), Hermite(qa + 9, dt), Hermite(qa + 18, dt) };
    const Hermite B[3] = { Hermite(qb, dt), Hermite(qb + 9, dt), Hermite(qb + 18, dt) };
    const Approach f { { &A[0], &A[1], &A[2] }, { &B[0], &B[1], &B[2] } };
    return root ? find_root(f, t, x1, tol) : f(t);
}

int orc_screen(const double* traj, int comps, int nobj, int nsteps, const unsigned* set,
    double* mind, int* mink)
{
    Partition p;
    if (!p.calc(set, nobj))
        return 1;
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (long q = 0; q < (long)p.numPairs; ++q) {
        unsigned a = 0, b = 0;
        p.pair((unsigned)q, a, b);
        double best = std::numeric_limits<double>::max();
        int bk = -1;
        for (int k = 0; k < nsteps; ++k) {
            const double* ta = traj + ((size_t)k * nobj + a) * comps;
            const double* tb = traj + ((size_t)k * nobj + b) * comps;
            double dd = dist_of(ta, tb);
            if (dd < best) {
                best = dd;
                bk = k;
            }
        }
        mind[q] = best;
        mink[q] = bk;
    }
    return 0;
}

// The same per-pair rule as orc_screen in a blocked loop order, for screens of 10^7-10^8 pairs
// over thousands of steps. orc_screen keeps, per pair, the first sample whose root is strictly
// below the running minimum; since sqrt is monotone and correctly rounded, its final distance is
// sqrt(min d2) and its step the first sample whose root equals that. So here the positions are
// transposed to [step][component][object] and tiles of 32 x 64 pairs keep min d2 over the whole
// time axis (a NaN is never taken, as under "dd < best"); the step is then found by a second
// scan, only for the pairs that need it (root below threshold, or every pair when
// threshold <= 0; the others get -1). Pairs without a sample below DBL_MAX get DBL_MAX and -1,
// as in orc_screen.
int orc_screen_blocked(const double* traj, int comps, int nobj, int nsteps, const unsigned* set,
    double threshold, double* mind, int* mink, int threads)
{
    Partition p;
    if (nobj <= 0 || !p.calc(set, nobj))
        return 1;
#ifdef _OPENMP
    const int nt = threads > 0 ? threads : omp_get_max_threads();
#else
    (void)threads;
#endif
    const size_t N = (size_t)nobj;
    std::vector<double> P((size_t)nsteps * 3 * N);
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nt)
#endif
    for (long k = 0; k < nsteps; ++k)
        for (size_t o = 0; o < N; ++o)
            for (int c = 0; c < 3; ++c)
                P[((size_t)k * 3 + c) * N + o] = traj[((size_t)k * N + o) * comps + c];
    constexpr int TA = 32, TB = 64;
    struct Tile {
        unsigned s, a0, na, b0, nb;
    };
    std::vector<Tile> tiles;
    for (unsigned s = 0; s + 1 < p.numSets; ++s) {
        const unsigned A0 = p.offsets[s], A1 = p.offsets[s + 1], B1 = (unsigned)nobj;
        for (unsigned a = A0; a < A1; a += TA)
            for (unsigned b = A1; b < B1; b += TB)
                tiles.push_back({ s, a, std::min<unsigned>(TA, A1 - a), b,
                    std::min<unsigned>(TB, B1 - b) });
    }
    const double dmax = std::numeric_limits<double>::max();
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
#endif
    for (long t = 0; t < (long)tiles.size(); ++t) {
        const Tile T = tiles[t];
        alignas(64) double m[TA][TB];
        for (unsigned i = 0; i < TA; ++i)
            for (unsigned j = 0; j < TB; ++j)
                m[i][j] = std::numeric_limits<double>::infinity();
        for (int k = 0; k < nsteps; ++k) {
            const double* px = &P[(size_t)k * 3 * N];
            const double* py = px + N;
            const double* pz = py + N;
            const double* bx = px + T.b0;
            const double* by = py + T.b0;
            const double* bz = pz + T.b0;
            for (unsigned i = 0; i < T.na; ++i) {
                const double ax = px[T.a0 + i], ay = py[T.a0 + i], az = pz[T.a0 + i];
                double* mi = m[i];
                for (unsigned j = 0; j < T.nb; ++j) {
                    const double dx = ax - bx[j], dy = ay - by[j], dz = az - bz[j];
                    const double d2 = std::fma(dz, dz, std::fma(dy, dy, dx * dx));
                    mi[j] = d2 < mi[j] ? d2 : mi[j];
                }
            }
        }
        const unsigned A0 = p.offsets[T.s], B0 = p.offsets[T.s + 1];
        for (unsigned i = 0; i < T.na; ++i)
            for (unsigned j = 0; j < T.nb; ++j) {
                const unsigned a = T.a0 + i, b = T.b0 + j;
                const size_t q = (size_t)p.mapoff[T.s] + (size_t)(a - A0) * p.higher[T.s] + (b - B0);
                const double root = std::sqrt(m[i][j]);
                if (!(root < dmax)) {
                    mind[q] = dmax;
                    mink[q] = -1;
                    continue;
                }
                mind[q] = root;
                mink[q] = -1;
                if (threshold > 0.0 && !(root < threshold))
                    continue;
                for (int k = 0; k < nsteps; ++k) {
                    const double* px = &P[(size_t)k * 3 * N];
                    const double dx = px[a] - px[b], dy = px[N + a] - px[N + b],
                                 dz = px[2 * N + a] - px[2 * N + b];
                    if (std::sqrt(std::fma(dz, dz, std::fma(dy, dy, dx * dx))) == root) {
                        mink[q] = k;
                        break;
                    }
                }
            }
    }
    return 0;
}

int orc_refine(const double* traj, int nobj, int nsteps, double h, unsigned a, unsigned b, int k,
    double mind, double* out)
{
    // out: distance, time, posA[3], posB[3]
    const double* ca = traj + ((size_t)k * nobj + a) * 9;
    const double* cb = traj + ((size_t)k * nobj + b) * 9;
    if (k == 0 || k == nsteps - 1) {
        out[0] = mind;
        out[1] = k == 0 ? 0.0 : (double)nsteps * h;
        for (int i = 0; i < 3; ++i) {
            out[2 + i] = ca[i];
            out[5 + i] = cb[i];
        }
        return 0;
    }
    auto mk = [&](unsigned o, int c) {
        double q[9];
        for (int j = 0; j < 3; ++j) {
            const double* t = traj + ((size_t)(k - 1 + j) * nobj + o) * 9;
            q[3 * j + 0] = t[c];
            q[3 * j + 1] = t[3 + c];
            q[3 * j + 2] = t[6 + c];
        }
        return Hermite(q, h);
    };
    Hermite Ax = mk(a, 0), Ay = mk(a, 1), Az = mk(a, 2);
    Hermite Bx = mk(b, 0), By = mk(b, 1), Bz = mk(b, 2);
    const Approach f { { &Ax, &Ay, &Az }, { &Bx, &By, &Bz } };
    double t = find_root(f, 0.0, h);
    out[1] = t + k * h;
    double dx = Ax.val(t) - Bx.val(t), dy = Ay.val(t) - By.val(t), dz = Az.val(t) - Bz.val(t);
    out[0] = std::sqrt(dx * dx + dy * dy + dz * dz);
    out[2] = Ax.val(t);
    out[3] = Ay.val(t);
    out[4] = Az.val(t);
    out[5] = Bx.val(t);
    out[6] = By.val(t);
    out[7] = Bz.val(t);
    return 0;
}


// Philox4x32-10: round r maps (x0, x1, x2, x3) with key (k0, k1) to
// (hi(B*x2) ^ x1 ^ k0, lo(B*x2), hi(A*x0) ^ x3 ^ k1, lo(A*x0)), A = 0xD2511F53, B = 0xCD9E8D57;
// the key grows by (0x9E3779B9, 0xBB67AE85) after every round.
void orc_philox4x32_10(const uint32_t* ctr4, const uint32_t* key2, uint32_t* out4)
{
    uint32_t x[4] = { ctr4[0], ctr4[1], ctr4[2], ctr4[3] };
    uint32_t k[2] = { key2[0], key2[1] };
    for (int r = 0; r < 10; ++r) {
        const uint64_t pa = static_cast<uint64_t>(0xD2511F53u) * x[0];
        const uint64_t pb = static_cast<uint64_t>(0xCD9E8D57u) * x[2];
        const uint32_t y0 = static_cast<uint32_t>(pb >> 32) ^ x[1] ^ k[0];
        const uint32_t y1 = static_cast<uint32_t>(pb & 0xffffffffu);
        const uint32_t y2 = static_cast<uint32_t>(pa >> 32) ^ x[3] ^ k[1];
        const uint32_t y3 = static_cast<uint32_t>(pa & 0xffffffffu);
        x[0] = y0;
        x[1] = y1;
        x[2] = y2;
        x[3] = y3;
        k[0] += 0x9E3779B9u;
        k[1] += 0xBB67AE85u;
    }
    for (int j = 0; j < 4; ++j)
        out4[j] = x[j];
}

void orc_perturbed(const double* r0, const double* v0, long n, double strength, uint64_t seed,
    double* r, double* v)
{
    const uint32_t key[2] = { static_cast<uint32_t>(seed & 0xffffffffu),
        static_cast<uint32_t>(seed >> 32) };
    // 53 bits of a word pair -> (0, 1), never 0 or 1
    auto unit = [](uint32_t hi, uint32_t lo) {
        const uint64_t m = (static_cast<uint64_t>(hi) << 21) | (lo >> 11);
        return (static_cast<double>(m) + 0.5) / 9007199254740992.0;
    };
    for (long i = 0; i < n; ++i) {
        const uint64_t idx = static_cast<uint64_t>(i);
        double z[4];
        for (uint32_t blk = 0; blk < 2; ++blk) {
            const uint32_t ctr[4] = { static_cast<uint32_t>(idx & 0xffffffffu),
                static_cast<uint32_t>(idx >> 32), blk, 0u };
            uint32_t w[4];
            orc_philox4x32_10(ctr, key, w);
            const double rad = std::sqrt(-2.0 * std::log(unit(w[0], w[1])));
            const double ang = (2.0 * kPi) * unit(w[2], w[3]);
            z[2 * blk] = rad * std::cos(ang);
            z[2 * blk + 1] = rad * std::sin(ang);
        }
        for (int c = 0; c < 3; ++c) {
            r[3 * i + c] = r0[c];
            v[3 * i + c] = v0[c] + z[c] * strength;
        }
    }
}


void orc_set_sources(int mask, double drag_bc)
{
    g_src_mask = mask;
    g_src_bc = drag_bc;
}

void orc_sun_moon(double mjd_tt, double* out6) { sun_moon((mjd_tt - kMjdJ2000) / 36525.0, out6); }

double orc_density(double h_m) { return density(h_m); }

void orc_sources_accel(int mask, double drag_bc, const double* tb6, const double* r, const double* v,
    double* a)
{
    add_sources(mask, drag_bc, tb6, r, v, a);
}

} // extern "C"
