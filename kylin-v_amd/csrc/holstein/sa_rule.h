// Host logic of the stochastic-adaptive one-site TDVP (SA-1TDVP).  Host-only C++.
//
// Follows the reference's pieces, with seeded draws in place of std::random_device:
//   quasi_gauss / inv_quasi_gauss   util/allocation.h:233-258 (with its pi = 3.1415926)
//   the enrichment count            holstein/svd.h:501-533
//   the SVS record and file         holstein/svd.h:358, :390-400; header holstein/tdvp.h:71-74
//   the width fit (the beta file)   bin/nls.py:1-58, read back by holstein/tdvp.h:343-358
// Draws: uniform [0, 1) from the top 53 bits of mix(mix(mix(seed ^ call) ^ stream) ^ index),
// mix = the SplitMix64 finaliser (Steele, Lea and Flood, OOPSLA 2014).  `call` numbers the
// enrichment SVDs of one sweep, stream 0 feeds the count rule (index = draw number), stream
// 1 + v the entries of the call's v-th new vector (index = position, mapped to [-1, 1)).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

namespace kv {

inline uint64_t sa_mix(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline double sa_u01(uint64_t seed, uint64_t call, uint64_t stream, uint64_t index) {
    uint64_t h = sa_mix(seed ^ call);
    h = sa_mix(h ^ stream);
    h = sa_mix(h ^ index);
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

// util/allocation.h:233-240: peak 1 at x = 0.
inline double quasi_gauss(double x, double bet) {
    const double pi_ref = 3.1415926;
    double r = std::exp(-0.25 * bet * (x * x)) * std::fabs(0.25 * x / bet);
    r += (0.5 / bet - 0.25 * (x * x)) * std::erfc(0.5 * std::sqrt(bet) * std::fabs(x)) * 0.5 *
         std::sqrt(pi_ref / bet);
    r *= std::exp(-0.25 * bet * (x * x)) * 4.0 * bet * std::sqrt(bet / pi_ref);
    return r;
}

// util/allocation.h:241-258: bisection on [0, 2] to width 1e-8 for quasi_gauss(x) = y;
// without a root in the bracket the interval closes on 2.
inline double inv_quasi_gauss(double y, double bet) {
    double up = 2.0, dn = 0.0, res = 0.0;
    while (up - dn > 1e-8) {
        res = 0.5 * (up + dn);
        if ((quasi_gauss(res, bet) - y) * (quasi_gauss(dn, bet) - y) < 0)
            up = res;
        else
            dn = res;
    }
    return res;
}

// Number of states the enrichment adds to a spectrum (descending, every sector's values):
// the tail is continued by 10^(+-r) s_last^2 / s_prev, r = inv_quasi_gauss(u_j, beta) for
// the j-th draw of stream 0, the sign alternating with the position, never rising, until
// a value falls to thres or the spectrum reaches maxdim; nothing unless 5 < n < maxdim.
inline int sa_extension_count(const std::vector<double>& desc, int maxdim, double thres, double beta,
                              uint64_t seed, uint64_t call) {
    const int n = (int)desc.size();
    if (!(n > 5 && n < maxdim)) return 0;
    std::vector<double> t{desc[n - 2], desc[n - 1]};
    int add = 0;
    for (int i = n, j = 0; i < maxdim; ++i, ++j) {
        const double rds = inv_quasi_gauss(sa_u01(seed, call, 0, (uint64_t)j), beta);
        const double last = t.back(), prev = t[t.size() - 2];
        const double s = (i + n) % 2 == 0 ? std::pow(10.0, rds) * (last * last) / prev
                                           : std::pow(10.0, -1.0 * rds) * (last * last) / prev;
        if (s <= thres) break;
        t.push_back(s);
        if (t.back() > t[t.size() - 2]) t.back() = t[t.size() - 2];
        add = (int)t.size() - 2;
    }
    return add;
}

// log10(s_i s_{i-2} / s_{i-1}^2), i >= 2, over a descending list of kept values.
inline std::vector<double> svs_record(const std::vector<double>& desc) {
    std::vector<double> r;
    for (size_t i = 2; i < desc.size(); ++i)
        r.push_back(std::log10(desc[i] * desc[i - 2] / (desc[i - 1] * desc[i - 1])));
    return r;
}

inline void svs_write_header(const std::string& path, int nsites) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) return;
    std::fprintf(f, "Total Site = %d\n", nsites);
    std::fclose(f);
}

// Appends one two-site SVD's record: the kept count, then one value per line.
inline void svs_append(const std::string& path, const std::vector<double>& desc) {
    FILE* f = std::fopen(path.c_str(), "a");
    if (!f) return;
    std::fprintf(f, "%d\n", (int)desc.size());
    for (double x : svs_record(desc)) std::fprintf(f, "%-20.11e\n", x);
    std::fclose(f);
}

// Per-bond value lists of an SVS file (bin/nls.py:19-48): records 1 .. N-1 open the lists
// of bonds 0 .. N-2, records N .. 2N-2 (the right-to-left sweep) join lists 2N-2-record;
// every later record joins list (record mod 2N-2), folded back from the right half, and
// where that list does not exist, the one before it.  Count lines are the lines shorter
// than 10 characters.  Returns false when the header is unreadable.
inline bool parse_svs(const std::string& text, std::vector<std::vector<double>>& lists) {
    std::vector<std::string> lines;
    size_t p = 0;
    while (p < text.size()) {
        size_t e = text.find('\n', p);
        if (e == std::string::npos) {
            lines.push_back(text.substr(p));
            break;
        }
        lines.push_back(text.substr(p, e - p + 1));
        p = e + 1;
    }
    size_t pos = 0;
    auto readline = [&]() -> std::string { return pos < lines.size() ? lines[pos++] : std::string(); };
    std::string ss = readline();
    int nsite = 0;
    if (std::sscanf(ss.c_str(), "Total Site = %d", &nsite) != 1 || nsite < 2) return false;
    const int crl = 2 * nsite - 2;
    int cur = 0;
    lists.clear();
    while (!ss.empty()) {
        ss = readline();
        if (ss.empty() || ss.size() >= 10) continue;
        cur += 1;
        const int nval = std::max(std::atoi(ss.c_str()) - 2, 0);
        if (cur <= crl) {
            std::vector<double> v;
            for (int i = 0; i < nval; ++i) v.push_back(std::strtod(readline().c_str(), nullptr));
            if (cur < nsite) {
                lists.push_back(v);
            } else {
                const int b = crl - cur;
                if (b < 0 || b >= (int)lists.size()) return false;
                lists[b].insert(lists[b].end(), v.begin(), v.end());
            }
        } else {
            int b = cur % crl;
            if (b > nsite - 1) b = crl - b;
            for (int i = 0; i < nval; ++i) {
                const double x = std::strtod(readline().c_str(), nullptr);
                const int t = b < (int)lists.size() ? b : b - 1;
                if (t < 0 || t >= (int)lists.size()) return false;
                lists[t].push_back(x);
            }
        }
    }
    return true;
}

// The fitted density (bin/nls.py:10-14); it integrates to one.
inline double sa_fit_model(double x, double bet) {
    const double pi = 3.14159265358979323846;
    double r = std::exp(-0.25 * bet * (x * x)) * std::fabs(x / 4.0 / bet);
    r += (1.0 / 2.0 / bet - (x * x) / 4.0) * std::erfc(std::sqrt(bet) * std::fabs(x) / 2.0) * 0.5 *
         std::sqrt(pi / bet);
    return r * std::exp(-bet * (x * x) / 4.0) * (bet * bet);
}

// numpy.histogram(v, 50): equal bins over [min, max] (widened by 0.5 each way when equal),
// the last bin closed; left edges and counts.
inline void sa_histogram(const std::vector<double>& v, std::vector<double>& left, std::vector<double>& count) {
    const int nb = 50;
    double lo = *std::min_element(v.begin(), v.end()), hi = *std::max_element(v.begin(), v.end());
    if (lo == hi) {
        lo -= 0.5;
        hi += 0.5;
    }
    std::vector<double> edge(nb + 1);
    const double step = (hi - lo) / nb;
    for (int i = 0; i < nb; ++i) edge[i] = (double)i * step + lo;
    edge[nb] = hi;
    count.assign(nb, 0.0);
    for (double a : v) {
        int k = (int)(((a - lo) / (hi - lo)) * nb);
        if (k == nb) k -= 1;
        if (a < edge[k]) k -= 1;
        else if (a >= edge[k + 1] && k != nb - 1) k += 1;
        count[k] += 1.0;
    }
    left.assign(edge.begin(), edge.begin() + nb);
}

// Least-squares width of one bond (bin/nls.py:51-55): the minimum of
// F(b) = sum_i (sa_fit_model(x_i, b) - y_i)^2 over b > 0 reached downhill from b = 1 on a
// grid of ln b with step 0.05 in [-30, 30], then golden-section refined to 1e-12 in ln b.
inline double sa_fit_width(const std::vector<double>& vals) {
    std::vector<double> x, c;
    sa_histogram(vals, x, c);
    std::vector<double> y(c.size());
    for (size_t i = 0; i < c.size(); ++i) y[i] = c[i] / (double)vals.size();
    auto F = [&](double t) {
        const double b = std::exp(t);
        double s = 0;
        for (size_t i = 0; i < x.size(); ++i) {
            const double d = sa_fit_model(x[i], b) - y[i];
            s += d * d;
        }
        return s;
    };
    const double h = 0.05, tmin = -30.0, tmax = 30.0;
    double t = 0.0, f = F(t);
    int dir = F(t + h) < f ? 1 : (F(t - h) < f ? -1 : 0);
    while (dir != 0) {
        const double tn = t + dir * h;
        if (tn < tmin || tn > tmax) break;
        const double fn = F(tn);
        if (!(fn < f)) break;
        t = tn;
        f = fn;
    }
    double a = t - h, b = t + h;
    const double gr = 0.5 * (std::sqrt(5.0) - 1.0);
    double c1 = b - gr * (b - a), c2 = a + gr * (b - a);
    double f1 = F(c1), f2 = F(c2);
    while (b - a > 1e-12) {
        if (f1 < f2) {
            b = c2;
            c2 = c1;
            f2 = f1;
            c1 = b - gr * (b - a);
            f1 = F(c1);
        } else {
            a = c1;
            c1 = c2;
            f1 = f2;
            c2 = a + gr * (b - a);
            f2 = F(c2);
        }
    }
    const double tb = 0.5 * (a + b);
    return std::exp(F(tb) <= f ? tb : t);
}

// Widths of every bond of an SVS file (1.0 for bonds with at most two values).
inline bool sa_fit_widths(const std::string& text, std::vector<double>& widths) {
    std::vector<std::vector<double>> lists;
    if (!parse_svs(text, lists)) return false;
    widths.clear();
    for (const auto& v : lists) widths.push_back(v.size() > 2 ? sa_fit_width(v) : 1.0);
    return true;
}

}  // namespace kv
