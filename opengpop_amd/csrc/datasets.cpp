#include "datasets.hpp"

#include <cmath>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace ogp {

namespace {
constexpr double kArcsPerRad = 3600.0 * 180.0 / 3.14159265358979324;
constexpr double kMjdNtpEpoch = 15020.0; // 1900-01-01

size_t tri(size_t n, size_t m) { return n * (n + 1) / 2 + m; }

std::ifstream open_or_throw(const std::string& path)
{
    std::ifstream f(path);
    if (!f.is_open())
        throw std::runtime_error("cannot open dataset file: " + path);
    return f;
}
} // namespace

GravityModel load_gravity(const std::string& path, int degree)
{
    std::ifstream f = open_or_throw(path);
    GravityModel g;
    std::string line;
    bool head_done = false;
    while (std::getline(f, line)) {
        if (line.find("end_of_head") != std::string::npos) {
            head_done = true;
            break;
        }
        std::istringstream is(line);
        std::string key;
        is >> key;
        if (key == "earth_gravity_constant")
            is >> g.gm;
        else if (key == "radius")
            is >> g.radius;
        else if (key == "max_degree")
            is >> g.file_max;
    }
    if (!head_done)
        throw std::runtime_error(path + ": invalid ICGEM file, 'end_of_head' not found");
    if (!(g.gm > 0) || !(g.radius > 0) || g.file_max < 0)
        throw std::runtime_error(path + ": missing earth_gravity_constant/radius/max_degree");
    if (degree > g.file_max)
        throw std::invalid_argument("requested degree " + std::to_string(degree)
            + " exceeds the file's max_degree " + std::to_string(g.file_max));
    g.L = degree < 0 ? g.file_max : degree;
    const int N = g.L + 1;
    g.C.assign(tri(N, 0), 0.0);
    g.S.assign(tri(N, 0), 0.0);
    bool first = true;
    while (std::getline(f, line)) {
        if (line.empty())
            continue;
        std::istringstream is(line);
        std::string key;
        is >> key;
        if (key.empty())
            continue;
        if (first) {
            if (key != "gfc")
                throw std::runtime_error(path + ": only static (gfc) models are accepted");
            first = false;
        }
        if (key != "gfc")
            continue;
        int n, m;
        double c, s;
        if (is >> n >> m >> c >> s)
            if (n >= 0 && n < N && m >= 0 && m <= n) {
                g.C[tri(n, m)] = c;
                g.S[tri(n, m)] = s;
            }
    }

    // Normalised derived-Legendre constants (Gottlieb 1993, eq. 7-8, 7-10, 7-15, 7-19).
    std::vector<double> pdiag(N, 0.0);
    pdiag[0] = 1.0;
    if (N > 1)
        pdiag[1] = std::sqrt(3.0);
    for (int i = 2; i < N; ++i)
        pdiag[i] = pdiag[i - 1] * std::sqrt((2.0 * i + 1.0) / (2.0 * i));
    g.diag_ratio.assign(N, 0.0);
    if (N > 1)
        g.diag_ratio[1] = std::sqrt(3.0);
    for (int i = 2; i < N; ++i)
        g.diag_ratio[i] = std::sqrt((2.0 * i + 1.0) / (2.0 * i));

    auto zeta = [](double n, double m) {
        return m == 0 ? std::sqrt(n * (n + 1.0) / 2.0) : std::sqrt((n - m) * (n + m + 1.0));
    };
    g.records.clear();
    g.records.reserve(kRecordLen * tri(N, 0));
    g.jac_ext.clear();
    g.jac_ext.reserve(kJacExtLen * tri(N, 0));
    std::vector<double> scale(N + 1); // c_n of the current column
    for (int m = 0; m <= g.L; ++m) {
        for (int n = m; n <= g.L; ++n) {
            double xi = 0;
            if (n <= m + 1) {
                scale[n] = 1.0;
                xi = n == m + 1 ? std::sqrt(2.0 * m + 3.0) : 0.0;
            } else {
                const double dn = n, dm = m;
                const double x = std::sqrt((2.0 * dn - 1.0) * (2.0 * dn + 1.0) / ((dn - dm) * (dn + dm)));
                const double eta = std::sqrt((dn + dm - 1.0) * (2.0 * dn + 1.0) * (dn - dm - 1.0)
                    / ((dn + dm) * (dn - dm) * (2.0 * dn - 3.0)));
                scale[n] = eta * scale[n - 2];
                xi = x * scale[n - 1] / scale[n];
            }
            const double cn = scale[n];
            double c = 0, s = 0, kc = 0, ks = 0, zc = 0, zs = 0;
            if (n >= 2) {
                c = g.C[tri(n, m)];
                s = g.S[tri(n, m)];
                kc = (1.0 + n + m) * c;
                ks = (1.0 + n + m) * s;
                if (m >= 1) {
                    double z = zeta(n, m - 1);
                    zc = z * g.C[tri(n, m - 1)];
                    zs = z * g.S[tri(n, m - 1)];
                }
            }
            const double rec[kRecordLen] = { xi, m * (cn * c), m * (cn * s), cn * kc, cn * ks,
                cn * zc, cn * zs };
            g.records.insert(g.records.end(), rec, rec + kRecordLen);
            double tc = 0, ts = 0;
            if (n >= 2 && m >= 2) {
                const double z2 = zeta(n, m - 2) * zeta(n, m - 1);
                tc = z2 * g.C[tri(n, m - 2)];
                ts = z2 * g.S[tri(n, m - 2)];
            }
            const double ext[kJacExtLen] = { (double)n, cn * tc, cn * ts, 0.0 };
            g.jac_ext.insert(g.jac_ext.end(), ext, ext + kJacExtLen);
        }
    }
    g.records_f.assign(g.records.begin(), g.records.end());
    return g;
}

namespace {

void load_nutation(const std::string& path, std::vector<NutationTerm>& j0,
    std::vector<NutationTerm>& j1)
{
    std::ifstream f = open_or_throw(path);
    std::string line;
    int block = -1;
    while (std::getline(f, line)) {
        if (line.find("j = 0") != std::string::npos) {
            block = 0;
            continue;
        }
        if (line.find("j = 1") != std::string::npos) {
            block = 1;
            continue;
        }
        std::istringstream is(line);
        int index;
        NutationTerm t;
        if (!(is >> index >> t.sin_uas >> t.cos_uas))
            continue;
        bool ok = true;
        for (int k = 0; k < 14 && ok; ++k)
            ok = static_cast<bool>(is >> t.mult[k]);
        if (!ok)
            throw std::runtime_error(path + ": malformed nutation row: " + line);
        if (block == 0)
            j0.push_back(t);
        else if (block == 1)
            j1.push_back(t);
    }
    if (j0.empty())
        throw std::runtime_error(path + ": no nutation terms found");
}

std::vector<EopRecord> load_eop(const std::string& path)
{
    std::ifstream f = open_or_throw(path);
    std::vector<EopRecord> out;
    std::string line;
    while (std::getline(f, line)) {
        size_t end = line.find_last_not_of(" \t\n\r\f\v");
        line.erase(end == std::string::npos ? 0 : end + 1);
        if (line.empty() || line[0] == '#' || line.size() < 68)
            continue;
        try {
            EopRecord r;
            r.mjd = std::stod(line.substr(7, 8));
            r.xp = std::stod(line.substr(18, 9)) / kArcsPerRad;
            r.yp = std::stod(line.substr(37, 9)) / kArcsPerRad;
            r.ut1_utc = std::stod(line.substr(58, 10));
            out.push_back(r);
        } catch (const std::exception&) {
            // rows without Bulletin A values (future dates) are skipped
        }
    }
    if (out.empty())
        throw std::runtime_error(path + ": no EOP rows found");
    return out;
}

void load_leap(const std::string& path, std::vector<double>& mjd, std::vector<double>& sec)
{
    std::ifstream f = open_or_throw(path);
    std::string line;
    while (std::getline(f, line)) {
        size_t p = line.find_first_not_of(" \t");
        if (p == std::string::npos || line[p] == '#')
            continue;
        std::istringstream is(line);
        long long ntp;
        double dtai;
        if (is >> ntp >> dtai) {
            mjd.push_back(kMjdNtpEpoch + static_cast<double>(ntp) / 86400.0);
            sec.push_back(dtai);
        }
    }
    if (mjd.empty())
        throw std::runtime_error(path + ": no leap seconds found");
}

} // namespace

FrameData load_frames(const std::string& dir)
{
    FrameData d;
    load_nutation(dir + "/IAU2000_R06_NutLon.txt", d.lon0, d.lon1);
    load_nutation(dir + "/IAU2000_R06_NutObl.txt", d.obl0, d.obl1);
    d.eop = load_eop(dir + "/IAU2000A_EOP.txt");
    load_leap(dir + "/leap-seconds.list", d.leap_mjd, d.leap_sec);
    return d;
}

} // namespace ogp
