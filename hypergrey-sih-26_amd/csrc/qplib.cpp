// QPLIB reader (host).
//
// A QPLIB file (QPLIB documentation, qplib.zib.de/doc.html) is a sequence of
// whitespace-separated words; '#' starts a comment that runs to the end of its
// line. The words follow a fixed order of sections:
//   name, three-letter type, objective sense;
//   n, m (type ...L only), the objective's quadratic part as one-based
//   lower-triangle triplets of Q in 1/2 x'Qx;
//   the linear objective and its constant; the constraint matrix (...L);
//   the infinity marker; constraint bounds (...L); variable bounds (...B, ...L);
//   three starting-point vectors; variable and constraint names.
// Every vector is written as a default value, a count and that many (index,
// value) pairs. Only continuous types with a quadratic or linear objective
// and at most linear constraints are read here: (C|D|Q) C (N|B|L).
//
// The parse fixes the model the solvers see, and so follows the reference
// reader (02-software/solver/src/qp.cpp:67-117) where its results depend on
// it: numbers are read as long double (the whole word, finite, no range
// error) and rounded once to double; n and m are at most 100,000 and the
// entry counts at most 5,000,000; a listed off-diagonal Q entry contributes
// half its value to each mirrored position (qp.cpp:84), a diagonal one its
// full value; |v| at or above the infinity marker reads as an infinite bound;
// Hessian triplets and each row's matrix entries keep their file order;
// duplicates, upper-triangle entries and words after the last section are
// errors (a later line may hold blanks and comments only).
//
// Structure: the text is cut into words in one pass (a line table kept for the
// trailing-content rule), then a cursor walks the sections, each a function of
// its own.
#include "hgqp/accelerated_qp.hpp"
#include "internal.hpp"

#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <iterator>
#include <limits>
#include <unordered_set>

namespace hgqp {
namespace {

constexpr size_t kMaxDimension = 100000;   // n and m (qp.cpp:75-76)
constexpr size_t kMaxEntries = 5000000;    // Hessian and matrix entry counts
constexpr double kInf = std::numeric_limits<double>::infinity();

bool blank(unsigned char ch) { return ch == ' ' || ch == '\t' || ch == '\n' || ch == '\v' || ch == '\f' || ch == '\r'; }

[[noreturn]] void fail(const std::string& what) { throw std::runtime_error("QPLIB: " + what); }

// The words of a text and where each ends.
class Lexed {
public:
    explicit Lexed(std::string text) : text_(std::move(text)) {
        size_t i = 0;
        const size_t len = text_.size();
        while (i < len) {
            const unsigned char ch = static_cast<unsigned char>(text_[i]);
            if (ch == '#') {  // comment: skip to the end of the line
                while (i < len && text_[i] != '\n') ++i;
            } else if (blank(ch)) {
                ++i;
            } else {
                const size_t from = i;
                while (i < len && !blank(static_cast<unsigned char>(text_[i])) && text_[i] != '#') ++i;
                spans_.push_back({from, i - from});
            }
        }
    }
    size_t size() const { return spans_.size(); }
    std::string word(size_t k) const { return text_.substr(spans_[k].first, spans_[k].second); }

    // After the last word that was read (k words consumed): nothing but blanks
    // and a comment may follow on its line, and every later line may hold only
    // spaces, tabs and carriage returns before an optional comment.
    void require_end(size_t consumed) const {
        if (consumed < spans_.size()) {
            // any further word is an error when it shares the line of the last
            // one read; a word on a later line fails the line rule below as well
            fail("unexpected words after the last section");
        }
        size_t i = consumed ? spans_[consumed - 1].first + spans_[consumed - 1].second : 0;
        // rest of the last word's line: comment and blanks only (no words remain)
        while (i < text_.size() && text_[i] != '\n') ++i;
        bool in_comment = false;
        for (; i < text_.size(); ++i) {
            const char ch = text_[i];
            if (ch == '\n') {
                in_comment = false;
            } else if (ch == '#') {
                in_comment = true;
            } else if (!in_comment && ch != ' ' && ch != '\t' && ch != '\r') {
                fail("unexpected content after the last section");
            }
        }
    }

private:
    std::string text_;
    std::vector<std::pair<size_t, size_t>> spans_;  // (first byte, length)
};

// Reads the words of a lexed text in order, with the format's value types.
class Cursor {
public:
    explicit Cursor(const Lexed& lx) : lx_(lx) {}
    size_t consumed() const { return next_; }

    std::string word() {
        if (next_ >= lx_.size()) fail("input ends before the model is complete");
        return lx_.word(next_++);
    }

    // A finite real number written as the whole word, read in long double.
    long double real() {
        const std::string w = word();
        errno = 0;
        char* stop = nullptr;
        const long double v = std::strtold(w.c_str(), &stop);
        if (stop == w.c_str() || errno == ERANGE || stop != w.c_str() + w.size() || !std::isfinite(v))
            fail("'" + w + "' is not a finite number");
        return v;
    }

    // A whole number in [0, limit].
    size_t whole(size_t limit) {
        const long double v = real();
        if (v < 0 || v > static_cast<long double>(limit) || std::floor(v) != v)
            fail("count out of range (at most " + std::to_string(limit) + ")");
        return static_cast<size_t>(v);
    }

    // A one-based position in [1, limit], returned zero-based.
    size_t position(size_t limit) {
        const size_t p = whole(limit);
        if (p == 0) fail("positions are one-based");
        return p - 1;
    }

    // A real rounded to double, infinite at or beyond `marker` in magnitude.
    double bounded(long double marker) {
        const long double v = real();
        if (std::fabs(v) >= marker) return v < 0 ? -kInf : kInf;
        if (std::fabs(v) > static_cast<long double>(std::numeric_limits<double>::max()))
            fail("value exceeds the double range");
        return static_cast<double>(v);
    }

    // A vector of length len: default value, count, then (position, value) pairs.
    std::vector<double> dense(size_t len, long double marker = std::numeric_limits<long double>::infinity()) {
        std::vector<double> out(len, bounded(marker));
        const size_t listed = whole(len);
        std::vector<char> set(len, 0);
        for (size_t k = 0; k < listed; ++k) {
            const size_t j = position(len);
            if (set[j]) fail("vector entry listed twice");
            set[j] = 1;
            out[j] = bounded(marker);
        }
        return out;
    }

private:
    const Lexed& lx_;
    size_t next_ = 0;
};

// ---- sections --------------------------------------------------------------

struct Header {
    char objective = 0, variables = 0, constraints = 0;  // the three type letters
    bool maximize = false;
};

Header read_header(Cursor& cur, LinearProblem& p) {
    p.name = cur.word();
    const std::string type = cur.word();
    const std::string sense = cur.word();
    Header h;
    if (type.size() == 3) {
        h.objective = type[0];
        h.variables = type[1];
        h.constraints = type[2];
    }
    const bool objective_ok = h.objective == 'C' || h.objective == 'D' || h.objective == 'Q';
    const bool constraints_ok = h.constraints == 'N' || h.constraints == 'B' || h.constraints == 'L';
    if (!objective_ok || h.variables != 'C' || !constraints_ok)
        fail("unsupported_qplib_type: continuous quadratic objective with linear/box constraints required");
    if (sense == "maximize") h.maximize = true;
    else if (sense != "minimize") fail("objective sense must be minimize or maximize");
    p.maximize = h.maximize;
    return h;
}

// Lower-triangle triplets (i >= j) of Q in 1/2 x'Qx, kept in file order; the
// model stores the full value of each mirrored position.
void read_quadratic(Cursor& cur, size_t n, std::vector<QuadraticTerm>& out) {
    const size_t listed = cur.whole(kMaxEntries);
    std::unordered_set<uint64_t> seen;
    seen.reserve(listed);
    out.reserve(listed);
    for (size_t k = 0; k < listed; ++k) {
        const size_t i = cur.position(n), j = cur.position(n);
        const double v = static_cast<double>(cur.real());
        if (i < j) fail("quadratic entry above the diagonal");
        if (!std::isfinite(v)) fail("quadratic entry is not finite in double");
        if (!seen.insert(uint64_t(i) * n + j).second) fail("quadratic entry listed twice");
        out.push_back({i, j, i == j ? v : 0.5 * v});
    }
}

// Constraint matrix triplets (row, column, value): each row keeps its entries
// in file order.
void read_matrix(Cursor& cur, size_t n, size_t m, LinearProblem& p) {
    const size_t listed = cur.whole(kMaxEntries);
    p.rows.assign(m, {});
    std::unordered_set<uint64_t> seen;
    seen.reserve(listed);
    for (size_t k = 0; k < listed; ++k) {
        const size_t i = cur.position(m), j = cur.position(n);
        const double v = static_cast<double>(cur.real());
        if (!std::isfinite(v)) fail("matrix entry is not finite in double");
        if (!seen.insert(uint64_t(i) * n + j).second) fail("matrix entry listed twice");
        p.rows[i].push_back({j, v});
    }
}

// (position, name) pairs; `keep` receives the names (variables), or null.
void read_names(Cursor& cur, size_t len, std::vector<std::string>* keep) {
    const size_t listed = cur.whole(len);
    std::vector<char> named(len, 0);
    for (size_t k = 0; k < listed; ++k) {
        const size_t j = cur.position(len);
        if (named[j]) fail("name listed twice");
        named[j] = 1;
        std::string name = cur.word();
        if (keep) (*keep)[j] = std::move(name);
    }
}

QuadraticProblem parse_qplib(const Lexed& lx) {
    Cursor cur(lx);
    QuadraticProblem q;
    LinearProblem& p = q.linear;
    const Header h = read_header(cur, p);
    const size_t n = cur.whole(kMaxDimension);
    if (n == 0) fail("the model has no variables");
    const size_t m = h.constraints == 'L' ? cur.whole(kMaxDimension) : 0;
    read_quadratic(cur, n, q.hessian);
    p.cost = cur.dense(n);
    p.offset = static_cast<double>(cur.real());
    if (m) read_matrix(cur, n, m, p);
    else p.rows.clear();
    const long double marker = cur.real();
    if (!(marker > 0)) fail("the infinity marker must be positive");
    if (m) {
        p.row_lower = cur.dense(m, marker);
        p.row_upper = cur.dense(m, marker);
    }
    if (h.constraints == 'N') {
        p.lower.assign(n, -kInf);
        p.upper.assign(n, kInf);
    } else {
        p.lower = cur.dense(n, marker);
        p.upper = cur.dense(n, marker);
    }
    // starting points (primal, constraint duals, bound duals): read and dropped
    (void)cur.dense(n);
    if (m) (void)cur.dense(m);
    (void)cur.dense(n);
    p.names.resize(n);
    for (size_t j = 0; j < n; ++j) p.names[j] = "x" + std::to_string(j + 1);
    p.integer.assign(n, false);
    read_names(cur, n, &p.names);
    read_names(cur, m, nullptr);
    lx.require_end(cur.consumed());
    return q;
}

}  // namespace

QuadraticProblem read_qplib(std::istream& input) {
    std::string text((std::istreambuf_iterator<char>(input)), std::istreambuf_iterator<char>());
    return parse_qplib(Lexed(std::move(text)));
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
        auto h = std::make_unique<hgqp_qplib>();
        try {
            h->q = parse_qplib(Lexed(std::string(text, length)));
        } catch (const std::runtime_error& e) {
            throw Error(HGQP_INVALID_ARGUMENT, e.what());
        }
        const size_t terms = h->q.hessian.size();
        h->hrow.resize(terms);
        h->hcol.resize(terms);
        h->hval.resize(terms);
        for (size_t k = 0; k < terms; ++k) {
            h->hrow[k] = int64_t(h->q.hessian[k].row);
            h->hcol[k] = int64_t(h->q.hessian[k].col);
            h->hval[k] = h->q.hessian[k].value;
        }
        const auto& rows = h->q.linear.rows;
        h->rstart.assign(1, 0);
        for (const auto& row : rows) {
            for (const auto& entry : row) {
                h->rcol.push_back(int64_t(entry.first));
                h->rval.push_back(entry.second);
            }
            h->rstart.push_back(int64_t(h->rcol.size()));
        }
        *out = h.release();
    });
}

extern "C" int hgqp_qplib_problem(const hgqp_qplib* model, hgqp_problem* view, int64_t* rows) {
    return guarded([&] {
        if (!model || !view) throw Error(HGQP_INVALID_ARGUMENT, "hgqp_qplib_problem: null argument");
        const LinearProblem& p = model->q.linear;
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
        const LinearProblem& p = model->q.linear;
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
