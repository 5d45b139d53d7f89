// opengpop command-line tool: the reference's `propagate` and `closest-approach` subcommands
// (exec/main.cpp:52-107, exec/runner.cu:11-119) on top of the C ABI of libogp.so.
//
// Input/output JSON follow the reference schema (main.cpp:28-40, runner.cuh:8-47); output is
// written like nlohmann::json::dump(4): keys in lexicographic order, four-space indentation,
// doubles in shortest round-trip form with ".0" on integral values.
//
// Extensions (all optional; defaults reproduce the reference): --step, --degree,
// --corrections, --precision fp64|split, --data-dir, --threshold (closest-approach: report only
// pairs whose sampled minimum distance is below the threshold, in metres), and the acceleration
// sources added to gravity: --sun, --moon, --drag C_D*A/m (m^2/kg); --timing prints the wall
// time of each phase (read and parse, context, compute, write) to stderr.
#include "ogp.h"

#include <array>
#include <charconv>
#include <chrono>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

// ---------------------------------------------------------------------------------------
// Minimal JSON value, parser and nlohmann-compatible writer
// ---------------------------------------------------------------------------------------
struct Json {
    enum Type { Null, Bool, Int, Float, String, Array, Object } type = Null;
    bool b = false;
    long long i = 0;
    double d = 0;
    std::string s;
    std::vector<Json> a;
    std::map<std::string, Json> o;

    static Json integer(long long v)
    {
        Json j;
        j.type = Int;
        j.i = v;
        return j;
    }
    static Json number(double v)
    {
        Json j;
        j.type = Float;
        j.d = v;
        return j;
    }
    static Json array() { Json j; j.type = Array; return j; }
    static Json object() { Json j; j.type = Object; return j; }

    const Json& at(const std::string& k) const
    {
        auto it = o.find(k);
        if (type != Object || it == o.end())
            throw std::runtime_error("missing key \"" + k + "\"");
        return it->second;
    }
    double num() const
    {
        if (type == Int)
            return (double)i;
        if (type == Float)
            return d;
        throw std::runtime_error("expected a number");
    }
    long long integer_value() const
    {
        if (type == Int)
            return i;
        if (type == Float && d == std::floor(d))
            return (long long)d;
        throw std::runtime_error("expected an integer");
    }
};

struct Parser {
    const std::string& t;
    size_t p = 0;
    explicit Parser(const std::string& text)
        : t(text)
    {
    }
    [[noreturn]] void fail(const char* what)
    {
        throw std::runtime_error(std::string("JSON parse error: ") + what + " at offset "
            + std::to_string(p));
    }
    void ws()
    {
        while (p < t.size() && (t[p] == ' ' || t[p] == '\n' || t[p] == '\r' || t[p] == '\t'))
            ++p;
    }
    bool lit(const char* w)
    {
        size_t n = std::strlen(w);
        if (t.compare(p, n, w) == 0) {
            p += n;
            return true;
        }
        return false;
    }
    Json value()
    {
        ws();
        if (p >= t.size())
            fail("unexpected end");
        char c = t[p];
        if (c == '{')
            return object();
        if (c == '[')
            return array();
        if (c == '"') {
            Json j;
            j.type = Json::String;
            j.s = str();
            return j;
        }
        if (lit("true")) {
            Json j;
            j.type = Json::Bool;
            j.b = true;
            return j;
        }
        if (lit("false")) {
            Json j;
            j.type = Json::Bool;
            return j;
        }
        if (lit("null"))
            return Json();
        return numeric();
    }
    std::string str()
    {
        std::string out;
        ++p;
        while (p < t.size() && t[p] != '"') {
            if (t[p] == '\\') {
                ++p;
                if (p >= t.size())
                    fail("bad escape");
                char e = t[p];
                out += e == 'n' ? '\n' : e == 't' ? '\t' : e == 'r' ? '\r' : e;
            } else {
                out += t[p];
            }
            ++p;
        }
        if (p >= t.size())
            fail("unterminated string");
        ++p;
        return out;
    }
    Json numeric()
    {
        size_t s = p;
        bool flt = false;
        if (t[p] == '-' || t[p] == '+')
            ++p;
        while (p < t.size()) {
            char c = t[p];
            if (c >= '0' && c <= '9')
                ++p;
            else if (c == '.' || c == 'e' || c == 'E' || c == '-' || c == '+') {
                flt = true;
                ++p;
            } else
                break;
        }
        std::string tok = t.substr(s, p - s);
        if (tok.empty() || tok == "-")
            fail("bad number");
        errno = 0;
        char* end = nullptr;
        if (!flt) {
            long long v = std::strtoll(tok.c_str(), &end, 10);
            if (errno == 0 && *end == 0)
                return Json::integer(v);
        }
        double d = std::strtod(tok.c_str(), &end);
        if (*end != 0)
            fail("bad number");
        return Json::number(d);
    }
    Json array()
    {
        Json j = Json::array();
        ++p;
        ws();
        if (p < t.size() && t[p] == ']') {
            ++p;
            return j;
        }
        for (;;) {
            j.a.push_back(value());
            ws();
            if (p < t.size() && t[p] == ',') {
                ++p;
                continue;
            }
            if (p < t.size() && t[p] == ']') {
                ++p;
                return j;
            }
            fail("expected , or ]");
        }
    }
    Json object()
    {
        Json j = Json::object();
        ++p;
        ws();
        if (p < t.size() && t[p] == '}') {
            ++p;
            return j;
        }
        for (;;) {
            ws();
            if (p >= t.size() || t[p] != '"')
                fail("expected key");
            std::string k = str();
            ws();
            if (p >= t.size() || t[p] != ':')
                fail("expected :");
            ++p;
            j.o[k] = value();
            ws();
            if (p < t.size() && t[p] == ',') {
                ++p;
                continue;
            }
            if (p < t.size() && t[p] == '}') {
                ++p;
                return j;
            }
            fail("expected , or }");
        }
    }
};

// Shortest decimal that round-trips, in the layout nlohmann::json uses for doubles.
std::string format_double(double v)
{
    if (!std::isfinite(v))
        return "null";
    if (v == 0)
        return std::signbit(v) ? "-0.0" : "0.0";
    // The smallest precision whose correctly rounded form reads back as v. No decimal shorter
    // than the shortest round-trip one (std::to_chars without a precision) reads back, so the
    // search starts at its digit count and normally ends there.
    char buf[64];
    std::to_chars_result r = std::to_chars(buf, buf + sizeof buf - 1, v, std::chars_format::scientific);
    int prec = 0;
    for (const char* c = buf; c < r.ptr && *c != 'e'; ++c)
        prec += *c >= '0' && *c <= '9';
    for (; prec <= 17; ++prec) {
        r = std::to_chars(buf, buf + sizeof buf - 1, v, std::chars_format::scientific, prec - 1);
        double back = 0.0;
        std::from_chars(buf, r.ptr, back);
        if (back == v)
            break;
    }
    *r.ptr = '\0';
    // buf = [-]d.ddde[+-]XX ; split into digits and decimal exponent
    std::string sbuf(buf);
    bool neg = sbuf[0] == '-';
    if (neg)
        sbuf = sbuf.substr(1);
    size_t epos = sbuf.find('e');
    std::string mant = sbuf.substr(0, epos);
    int exp10 = std::atoi(sbuf.c_str() + epos + 1);
    std::string digits;
    for (char c : mant)
        if (c != '.')
            digits += c;
    while (digits.size() > 1 && digits.back() == '0')
        digits.pop_back();
    const int k = (int)digits.size();
    const int n = exp10 + 1; // decimal point position relative to the digits
    std::string out;
    if (k <= n && n <= 15) {
        out = digits + std::string(n - k, '0') + ".0";
    } else if (0 < n && n <= 15) {
        out = digits.substr(0, n) + "." + digits.substr(n);
    } else if (-4 < n && n <= 0) {
        out = "0." + std::string(-n, '0') + digits;
    } else {
        out = digits.substr(0, 1);
        if (k > 1)
            out += "." + digits.substr(1);
        int e = n - 1;
        char eb[16];
        std::snprintf(eb, sizeof eb, "e%c%02d", e < 0 ? '-' : '+', std::abs(e));
        out += eb;
    }
    return (neg ? "-" : "") + out;
}

void dump(const Json& j, std::ostream& os, int indent)
{
    const std::string pad(indent, ' ');
    const std::string padin(indent + 4, ' ');
    switch (j.type) {
    case Json::Null: os << "null"; break;
    case Json::Bool: os << (j.b ? "true" : "false"); break;
    case Json::Int: os << j.i; break;
    case Json::Float: os << format_double(j.d); break;
    case Json::String: os << '"' << j.s << '"'; break;
    case Json::Array:
        if (j.a.empty()) {
            os << "[]";
            break;
        }
        os << "[\n";
        for (size_t q = 0; q < j.a.size(); ++q) {
            os << padin;
            dump(j.a[q], os, indent + 4);
            os << (q + 1 < j.a.size() ? ",\n" : "\n");
        }
        os << pad << "]";
        break;
    case Json::Object: {
        if (j.o.empty()) {
            os << "{}";
            break;
        }
        os << "{\n";
        size_t q = 0;
        for (const auto& kv : j.o) {
            os << padin << '"' << kv.first << "\": ";
            dump(kv.second, os, indent + 4);
            os << (++q < j.o.size() ? ",\n" : "\n");
        }
        os << pad << "}";
        break;
    }
    }
}

// ---------------------------------------------------------------------------------------
struct Options {
    std::string cmd, input, output, data_dir = "data", precision = "fp64";
    double step = 30.0; // IdealStepSize, runner.cuh:6
    int degree = -1, corrections = 7;
    double threshold = 0.0;
    int sources = 0;      // OGP_SOURCE_* bits
    double drag_bc = 0.0; // C_D A / m with --drag
    bool timing = false;  // --timing: phase wall times on stderr
};

[[noreturn]] void usage(const char* msg)
{
    if (msg)
        std::cerr << "error: " << msg << "\n";
    std::cerr << "usage: opengpop {propagate|closest-approach} -i INPUT [-o OUTPUT]\n"
                 "       [--step S] [--degree L] [--corrections K] [--precision fp64|split]\n"
                 "       [--data-dir DIR] [--threshold METRES] [--sun] [--moon] [--drag CDA_M]\n"
                 "       [--timing]\n";
    std::exit(2);
}

std::vector<std::array<double, 3>> vec3s(const Json& j, const char* what)
{
    if (j.type != Json::Array)
        throw std::runtime_error(std::string(what) + " must be an array of [x, y, z]");
    std::vector<std::array<double, 3>> out;
    for (const Json& e : j.a) {
        if (e.type != Json::Array || e.a.size() != 3)
            throw std::runtime_error(std::string(what) + " entries must have 3 numbers");
        out.push_back({ e.a[0].num(), e.a[1].num(), e.a[2].num() });
    }
    return out;
}

Json vec3_json(const std::vector<std::array<double, 3>>& v)
{
    Json a = Json::array();
    for (const auto& x : v) {
        Json e = Json::array();
        for (double c : x)
            e.a.push_back(Json::number(c));
        a.a.push_back(e);
    }
    return a;
}

struct Common {
    Json start;
    long long seconds = 0;
    std::vector<std::array<double, 3>> r, v;
    int ymdhm[5];
    double sec;
    int64_t nsteps;
    double h;
};

Common read_common(const Json& in, const Options& opt)
{
    Common c;
    c.start = in.at("propagationStart");
    const char* keys[5] = { "Year", "Month", "Day", "Hour", "Min" };
    for (int q = 0; q < 5; ++q)
        c.ymdhm[q] = (int)c.start.at(keys[q]).integer_value();
    c.sec = c.start.at("Sec").num();
    // the reference's TimeData.Sec is a double: always written back as a double
    c.start.o["Sec"] = Json::number(c.sec);
    c.seconds = in.at("propagationSeconds").integer_value();
    c.r = vec3s(in.at("positions"), "positions");
    c.v = vec3s(in.at("velocities"), "velocities");
    if (c.r.size() != c.v.size())
        throw std::runtime_error("positions and velocities differ in length");
    // runner.cu:13-14: n = floor(T / 30), h = T / n
    c.nsteps = (int64_t)(c.seconds / (long long)opt.step);
    if (opt.step != std::floor(opt.step))
        c.nsteps = (int64_t)std::floor((double)c.seconds / opt.step);
    if (c.nsteps < 1)
        throw std::runtime_error("propagationSeconds is shorter than one step");
    c.h = (double)c.seconds / (double)c.nsteps;
    return c;
}

ogp_context* make_context(const Options& opt, const Common& c)
{
    ogp_config cfg;
    ogp_default_config(&cfg);
    cfg.data_dir = opt.data_dir.c_str();
    cfg.degree = opt.degree;
    cfg.n_corrections = opt.corrections;
    cfg.precision = opt.precision == "split" ? OGP_PRECISION_SPLIT : OGP_PRECISION_FP64;
    cfg.year = c.ymdhm[0];
    cfg.month = c.ymdhm[1];
    cfg.day = c.ymdhm[2];
    cfg.hour = c.ymdhm[3];
    cfg.minute = c.ymdhm[4];
    cfg.second = c.sec;
    ogp_context* ctx = nullptr;
    if (ogp_create(&cfg, &ctx) != OGP_OK)
        throw std::runtime_error(ogp_last_error(nullptr));
    if (opt.sources) {
        ogp_sources s {};
        s.mask = opt.sources;
        s.drag_bc = opt.drag_bc;
        if (ogp_set_sources(ctx, &s) != OGP_OK) {
            const std::string msg = ogp_last_error(ctx);
            ogp_destroy(ctx);
            throw std::runtime_error(msg);
        }
    }
    return ctx;
}

void check(ogp_context* ctx, int rc)
{
    if (rc != OGP_OK)
        throw std::runtime_error(ogp_last_error(ctx));
}

int run(const Options& opt)
{
    using Clock = std::chrono::steady_clock;
    Clock::time_point tp = Clock::now();
    auto phase = [&](const char* name) {
        const Clock::time_point now = Clock::now();
        if (opt.timing)
            std::cerr << "timing: " << name << " "
                      << std::chrono::duration<double>(now - tp).count() << " s\n";
        tp = now;
    };
    std::ifstream f(opt.input);
    if (!f.is_open())
        throw std::runtime_error("cannot open input " + opt.input);
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string text = ss.str();
    Parser ps(text);
    Json in = ps.value();
    Common c = read_common(in, opt);
    const int64_t n = (int64_t)c.r.size();
    phase("read and parse");
    std::unique_ptr<ogp_context, void (*)(ogp_context*)> ctx(make_context(opt, c), ogp_destroy);
    phase("context (device, data sets, tables)");
    Json out = Json::object();
    out.o["propagationStart"] = c.start;
    out.o["propagationSeconds"] = Json::integer(c.seconds);
    out.o["positionsStart"] = vec3_json(c.r);
    out.o["velocitiesStart"] = vec3_json(c.v);
    if (opt.cmd == "propagate") {
        std::cout << "Running Propagation...\n";
        std::vector<std::array<double, 3>> r(n), v(n);
        std::vector<double> h(n, c.h);
        std::vector<int64_t> st(n, c.nsteps);
        check(ctx.get(), ogp_propagate(ctx.get(), &c.r[0][0], &c.v[0][0], n, h.data(), st.data(),
                             &r[0][0], &v[0][0]));
        out.o["positionsEnd"] = vec3_json(r);
        out.o["velocitiesEnd"] = vec3_json(v);
    } else {
        std::cout << "Running Closest Approach Calculation...\n";
        std::vector<uint32_t> sets;
        const Json& sj = in.at("setIndices");
        if (sj.type != Json::Array)
            throw std::runtime_error("setIndices must be an array");
        for (const Json& e : sj.a)
            sets.push_back((uint32_t)e.integer_value());
        if ((int64_t)sets.size() != n)
            throw std::runtime_error("setIndices and positions differ in length");
        check(ctx.get(), ogp_conj_init(ctx.get(), &c.r[0][0], &c.v[0][0], sets.data(), n, c.h));
        check(ctx.get(), ogp_conj_propagate(ctx.get(), c.nsteps));
        int64_t cnt = 0;
        check(ctx.get(), ogp_conj_find(ctx.get(), opt.threshold, &cnt));
        std::vector<uint32_t> a(cnt), b(cnt);
        std::vector<int32_t> k(cnt);
        std::vector<double> d(cnt), t(cnt), pa(3 * cnt), pb(3 * cnt);
        check(ctx.get(), ogp_conj_results(ctx.get(), a.data(), b.data(), k.data(), d.data(),
                             t.data(), pa.data(), pb.data()));
        Json recs = Json::array();
        for (int64_t q = 0; q < cnt; ++q) {
            Json r = Json::object();
            r.o["objectIndex1"] = Json::integer(a[q]);
            r.o["objectIndex2"] = Json::integer(b[q]);
            r.o["distance"] = Json::number(d[q]);
            r.o["time"] = Json::number(t[q]);
            const char* an[3] = { "positionAx", "positionAy", "positionAz" };
            const char* bn[3] = { "positionBx", "positionBy", "positionBz" };
            for (int cc = 0; cc < 3; ++cc) {
                r.o[an[cc]] = Json::number(pa[3 * q + cc]);
                r.o[bn[cc]] = Json::number(pb[3 * q + cc]);
            }
            recs.a.push_back(r);
        }
        out.o["closestApproaches"] = recs;
    }
    phase("compute");
    std::ofstream o(opt.output);
    if (!o.is_open())
        throw std::runtime_error("cannot write " + opt.output);
    dump(out, o, 0);
    o.close();
    phase("write");
    std::cout << "Saved to: " << opt.output << "\n";
    return 0;
}

} // namespace

int main(int argc, char** argv)
{
    if (argc < 2)
        usage(nullptr);
    Options opt;
    opt.cmd = argv[1];
    if (opt.cmd == "--format-double") { // used by the CPU tests of the JSON writer
        for (int q = 2; q < argc; ++q)
            std::cout << format_double(std::strtod(argv[q], nullptr)) << "\n";
        return 0;
    }
    if (opt.cmd != "propagate" && opt.cmd != "closest-approach")
        usage("a subcommand (propagate or closest-approach) is required");
    opt.output = opt.cmd == "propagate" ? "propagation_out.json" : "closest_approach_out.json";
    for (int q = 2; q < argc; ++q) {
        std::string a = argv[q];
        auto next = [&]() -> std::string {
            if (q + 1 >= argc)
                usage(("missing value for " + a).c_str());
            return argv[++q];
        };
        if (a == "-i" || a == "--input")
            opt.input = next();
        else if (a == "-o" || a == "--output")
            opt.output = next();
        else if (a == "--step")
            opt.step = std::atof(next().c_str());
        else if (a == "--degree")
            opt.degree = std::atoi(next().c_str());
        else if (a == "--corrections")
            opt.corrections = std::atoi(next().c_str());
        else if (a == "--precision")
            opt.precision = next();
        else if (a == "--data-dir")
            opt.data_dir = next();
        else if (a == "--threshold")
            opt.threshold = std::atof(next().c_str());
        else if (a == "--timing")
            opt.timing = true;
        else if (a == "--sun")
            opt.sources |= OGP_SOURCE_SUN;
        else if (a == "--moon")
            opt.sources |= OGP_SOURCE_MOON;
        else if (a == "--drag") {
            opt.sources |= OGP_SOURCE_DRAG;
            opt.drag_bc = std::atof(next().c_str());
        } else
            usage(("unknown option " + a).c_str());
    }
    if (opt.input.empty())
        usage("--input is required");
    if (!(opt.step > 0))
        usage("--step must be positive");
    if (opt.precision != "fp64" && opt.precision != "split")
        usage("--precision must be fp64 or split");
    if ((opt.sources & OGP_SOURCE_DRAG) && !(opt.drag_bc >= 0))
        usage("--drag needs C_D A / m >= 0 (m^2/kg)");
    {
        std::ifstream probe(opt.input);
        if (!probe.is_open())
            usage(("input file does not exist: " + opt.input).c_str());
    }
    try {
        return run(opt);
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
}
