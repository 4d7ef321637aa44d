// QPLIB reader (host): the format as the reference reads it
// (02-software/solver/src/qp.cpp:12-117): whitespace tokens, '#' comments,
// one-based indices, default-plus-sparse vectors with an infinity marker, and
// Hessian entries listed for the lower triangle of Q in 1/2 x'Qx, so an
// off-diagonal entry contributes half its value to each mirrored position
// (qp.cpp:84). Continuous types (C|D|Q) x C x (N|B|L) only.
#include "hgqp/accelerated_qp.hpp"
#include "internal.hpp"

#include <cmath>
#include <cstring>
#include <fstream>
#include <set>
#include <sstream>

namespace hgqp {
namespace {

struct Reader {
    std::istream& in;
    explicit Reader(std::istream& stream) : in(stream) {}
    std::vector<std::string> pending;
    std::size_t at = 0;
    std::string token() {
        while (at == pending.size()) {
            std::string line;
            if (!std::getline(in, line)) throw std::runtime_error("Truncated QPLIB input");
            const auto hash = line.find('#');
            if (hash != std::string::npos) line.resize(hash);
            std::istringstream s(line);
            pending.clear();
            at = 0;
            std::string v;
            while (s >> v) pending.push_back(v);
        }
        return pending[at++];
    }
    long double number() {
        const std::string v = token();
        std::size_t used = 0;
        long double x = std::stold(v, &used);
        if (used != v.size() || !std::isfinite(x)) throw std::runtime_error("Invalid QPLIB number");
        return x;
    }
    std::size_t count(std::size_t maximum = 5000000) {
        const long double v = number();
        if (v < 0 || v > maximum || std::floor(v) != v) throw std::runtime_error("Invalid or oversized QPLIB count");
        return static_cast<std::size_t>(v);
    }
    std::size_t index(std::size_t n) {
        const auto i = count(n);
        if (i == 0) throw std::runtime_error("QPLIB indices are one-based");
        return i - 1;
    }
    std::vector<double> vector(std::size_t n, long double infinity = std::numeric_limits<long double>::infinity()) {
        auto convert = [&](long double v) {
            if (std::abs(v) >= infinity)
                return v < 0 ? -std::numeric_limits<double>::infinity() : std::numeric_limits<double>::infinity();
            if (std::abs(v) > std::numeric_limits<double>::max()) throw std::runtime_error("QPLIB value overflow");
            return static_cast<double>(v);
        };
        const double fill = convert(number());
        std::vector<double> result(n, fill);
        const auto entries = count(n);
        std::set<std::size_t> seen;
        for (std::size_t k = 0; k < entries; ++k) {
            const auto j = index(n);
            if (!seen.insert(j).second) throw std::runtime_error("Duplicate QPLIB vector entry");
            result[j] = convert(number());
        }
        return result;
    }
};

}  // namespace

QuadraticProblem read_qplib(std::istream& input) {
    Reader r(input);
    QuadraticProblem q;
    auto& p = q.linear;
    p.name = r.token();
    const std::string type = r.token(), sense = r.token();
    if (type.size() != 3 || type[1] != 'C' || std::string("CDQ").find(type[0]) == std::string::npos ||
        std::string("NBL").find(type[2]) == std::string::npos)
        throw std::runtime_error(
            "unsupported_qplib_type: continuous quadratic objective with linear/box constraints required");
    if (sense != "minimize" && sense != "maximize") throw std::runtime_error("Invalid QPLIB objective sense");
    p.maximize = sense == "maximize";
    const auto n = r.count(100000);
    if (!n) throw std::runtime_error("Empty QPLIB model");
    const auto m = type[2] == 'L' ? r.count(100000) : 0;
    const auto nz = r.count();
    std::set<std::pair<std::size_t, std::size_t> > seen;
    for (std::size_t k = 0; k < nz; ++k) {
        const auto i = r.index(n), j = r.index(n);
        const double v = static_cast<double>(r.number());
        if (i < j || !seen.insert({i, j}).second || !std::isfinite(v)) throw std::runtime_error("Invalid QPLIB Hessian entry");
        q.hessian.push_back({i, j, i == j ? v : 0.5 * v});
    }
    p.cost = r.vector(n);
    p.offset = static_cast<double>(r.number());
    p.rows.resize(m);
    if (m) {
        const auto count = r.count();
        std::set<std::pair<std::size_t, std::size_t> > entries;
        for (std::size_t k = 0; k < count; ++k) {
            const auto i = r.index(m), j = r.index(n);
            const double v = static_cast<double>(r.number());
            if (!entries.insert({i, j}).second || !std::isfinite(v)) throw std::runtime_error("Invalid QPLIB matrix entry");
            p.rows[i].push_back({j, v});
        }
    }
    const long double infinity = r.number();
    if (infinity <= 0) throw std::runtime_error("Invalid infinity marker");
    if (m) {
        p.row_lower = r.vector(m, infinity);
        p.row_upper = r.vector(m, infinity);
    }
    if (type[2] != 'N') {
        p.lower = r.vector(n, infinity);
        p.upper = r.vector(n, infinity);
    } else {
        p.lower.assign(n, -std::numeric_limits<double>::infinity());
        p.upper.assign(n, std::numeric_limits<double>::infinity());
    }
    // starting points are parsed and not used
    r.vector(n);
    if (m) r.vector(m);
    r.vector(n);
    p.names.resize(n);
    p.integer.assign(n, false);
    for (std::size_t j = 0; j < n; ++j) p.names[j] = "x" + std::to_string(j + 1);
    auto names = r.count(n);
    std::set<std::size_t> named;
    for (std::size_t k = 0; k < names; ++k) {
        const auto j = r.index(n);
        if (!named.insert(j).second) throw std::runtime_error("Duplicate variable name entry");
        p.names[j] = r.token();
    }
    names = r.count(m);
    named.clear();
    for (std::size_t k = 0; k < names; ++k) {
        const auto i = r.index(m);
        if (!named.insert(i).second) throw std::runtime_error("Duplicate row name entry");
        r.token();
    }
    if (r.at < r.pending.size()) throw std::runtime_error("Unexpected trailing QPLIB tokens");
    std::string line;
    while (std::getline(input, line)) {
        line = line.substr(0, line.find('#'));
        if (line.find_first_not_of(" \t\r\n") != std::string::npos)
            throw std::runtime_error("Unexpected trailing QPLIB content");
    }
    return q;
}

}  // namespace hgqp

using namespace hgqp;

// A parsed model kept alive for hgqp_qplib_problem's borrowed pointers.
struct hgqp_qplib {
    QuadraticProblem q;
    std::vector<int64_t> hrow, hcol;
    std::vector<double> hval;
    std::vector<int64_t> rstart, rcol;  // the linear rows as CSR
    std::vector<double> rval;
};

extern "C" int hgqp_qplib_parse(const char* text, size_t length, hgqp_qplib** out) {
    return guarded([&] {
        if (!text || !out) throw Error(HGQP_INVALID_ARGUMENT, "hgqp_qplib_parse: null argument");
        *out = nullptr;
        std::istringstream in(std::string(text, length));
        auto h = std::make_unique<hgqp_qplib>();
        try {
            h->q = read_qplib(in);
        } catch (const std::runtime_error& e) {
            throw Error(HGQP_INVALID_ARGUMENT, e.what());
        } catch (const std::logic_error& e) {  // std::stold on a malformed token
            throw Error(HGQP_INVALID_ARGUMENT, std::string("Invalid QPLIB number: ") + e.what());
        }
        for (const auto& t : h->q.hessian) {
            h->hrow.push_back(int64_t(t.row));
            h->hcol.push_back(int64_t(t.col));
            h->hval.push_back(t.value);
        }
        h->rstart.push_back(0);
        for (const auto& row : h->q.linear.rows) {
            for (const auto& a : row) {
                h->rcol.push_back(int64_t(a.first));
                h->rval.push_back(a.second);
            }
            h->rstart.push_back(int64_t(h->rcol.size()));
        }
        *out = h.release();
    });
}

extern "C" int hgqp_qplib_problem(const hgqp_qplib* model, hgqp_problem* view, int64_t* rows) {
    return guarded([&] {
        if (!model || !view) throw Error(HGQP_INVALID_ARGUMENT, "hgqp_qplib_problem: null argument");
        const auto& p = model->q.linear;
        std::memset(view, 0, sizeof(*view));
        view->n = int64_t(p.cost.size());
        view->nterms = int64_t(model->hval.size());
        view->hrow = model->hrow.data();
        view->hcol = model->hcol.data();
        view->hval = model->hval.data();
        view->cost = p.cost.data();
        view->lower = p.lower.data();
        view->upper = p.upper.data();
        view->offset = p.offset;
        view->maximize = p.maximize ? 1 : 0;
        if (rows) *rows = int64_t(p.rows.size());
    });
}

extern "C" int hgqp_qplib_rows(const hgqp_qplib* model, hgqp_rows* rows) {
    return guarded([&] {
        if (!model || !rows) throw Error(HGQP_INVALID_ARGUMENT, "hgqp_qplib_rows: null argument");
        const auto& p = model->q.linear;
        rows->m = int64_t(p.rows.size());
        rows->start = model->rstart.data();
        rows->col = model->rcol.data();
        rows->val = model->rval.data();
        rows->row_lower = p.row_lower.data();
        rows->row_upper = p.row_upper.data();
    });
}

extern "C" int hgqp_qplib_free(hgqp_qplib* model) {
    return guarded([&] { delete model; });
}
