// spz_qasm.cpp -- OpenQASM 2.0 loader: program text -> spz_op list
// (openqasm.rs:13-171).  Host code only; it needs no device.
//
// The reference lexes and parses with the `qasm` crate (1.0) and evaluates
// the angle of `cp` with `evalexpr` (11.3) in a context where `pi` is PI.
// Its build_circuit accepts `qreg` declarations and applications of
// h x y z rx ry rz u cp cx to indexed qubits; every other statement reaches a
// todo!() (a panic).  This loader accepts the same programs and produces the
// ops the reference's builders would add (QuantumCircuit::h ... ::cx,
// circuit.rs:215-445): single-qubit gates apply to each listed qubit in
// order, `cp(theta) c, t` and `cx c, t` take the control first.
//
// Differences, all on programs the reference cannot load:
//  * angles of rx ry rz u may be expressions as for cp (the reference parses
//    them as float literals only and panics on anything else);
//  * `include "qelib1.inc";` is accepted and skipped (the reference's
//    pre-processor would splice the file in, whose gate definitions its
//    builder then rejects);
//  * a whole register as a gate argument is an error (the reference skips
//    it silently for single-qubit gates and panics for cp and cx);
//  * registers are laid out in declaration order (the reference iterates a
//    HashMap, so with several registers its qubit order is not fixed);
//  * errors come back as a status and a message with the line number.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/spinoza_amd.h"

namespace spz {
int fail(int code, const std::string &msg);
}

namespace {

const double kPi = 3.14159265358979323846264338327950288;  // consts.rs PI (f64)

struct Reg {
    std::string name;
    int size, shift;
};

struct Parser {
    const char *p, *end;
    int line = 1;
    std::string err;

    bool at_end() { skip_ws(); return p >= end; }

    void skip_ws()
    {
        while (p < end) {
            if (*p == '\n') { line++; p++; }
            else if (*p == ' ' || *p == '\t' || *p == '\r') p++;
            else if (*p == '/' && p + 1 < end && p[1] == '/') {
                while (p < end && *p != '\n') p++;
            } else break;
        }
    }

    bool error(const std::string &m)
    {
        if (err.empty()) err = "line " + std::to_string(line) + ": " + m;
        return false;
    }

    bool expect(char c)
    {
        skip_ws();
        if (p < end && *p == c) { p++; return true; }
        return error(std::string("expected '") + c + "'");
    }

    bool peek(char c) { skip_ws(); return p < end && *p == c; }

    bool ident(std::string &out)
    {
        skip_ws();
        const char *s = p;
        if (p < end && (isalpha((unsigned char)*p) || *p == '_')) {
            p++;
            while (p < end && (isalnum((unsigned char)*p) || *p == '_')) p++;
        }
        if (p == s) return error("expected an identifier");
        out.assign(s, p);
        return true;
    }

    bool integer(long &out)
    {
        skip_ws();
        const char *s = p;
        while (p < end && isdigit((unsigned char)*p)) p++;
        if (p == s) return error("expected an integer");
        out = strtol(std::string(s, p).c_str(), nullptr, 10);
        return true;
    }

    // Angle expressions: numbers, pi, unary minus, + - * / ^ and parentheses,
    // with the usual precedence (^ binds tightest and to the right).
    bool expr(double &v)
    {
        if (!term(v)) return false;
        for (;;) {
            skip_ws();
            if (p < end && (*p == '+' || *p == '-')) {
                char op = *p++;
                double r;
                if (!term(r)) return false;
                v = op == '+' ? v + r : v - r;
            } else {
                return true;
            }
        }
    }

    bool term(double &v)
    {
        if (!unary(v)) return false;
        for (;;) {
            skip_ws();
            if (p < end && (*p == '*' || *p == '/')) {
                char op = *p++;
                double r;
                if (!unary(r)) return false;
                v = op == '*' ? v * r : v / r;
            } else {
                return true;
            }
        }
    }

    bool unary(double &v)
    {
        skip_ws();
        if (p < end && *p == '-') {
            p++;
            if (!unary(v)) return false;
            v = -v;
            return true;
        }
        if (p < end && *p == '+') {
            p++;
            return unary(v);
        }
        return power(v);
    }

    bool power(double &v)
    {
        if (!atom(v)) return false;
        skip_ws();
        if (p < end && *p == '^') {
            p++;
            double r;
            if (!unary(r)) return false;
            v = pow(v, r);
        }
        return true;
    }

    bool atom(double &v)
    {
        skip_ws();
        if (p < end && *p == '(') {
            p++;
            if (!expr(v)) return false;
            return expect(')');
        }
        if (p < end && (isdigit((unsigned char)*p) || *p == '.')) {
            const char *s = p;
            while (p < end && (isdigit((unsigned char)*p) || *p == '.')) p++;
            if (p < end && (*p == 'e' || *p == 'E')) {
                const char *q = p + 1;
                if (q < end && (*q == '+' || *q == '-')) q++;
                if (q < end && isdigit((unsigned char)*q)) {
                    p = q;
                    while (p < end && isdigit((unsigned char)*p)) p++;
                }
            }
            // correctly rounded decimal -> f64, as Rust's str::parse::<f64>
            v = strtod(std::string(s, p).c_str(), nullptr);
            return true;
        }
        std::string id;
        if (!ident(id)) return false;
        if (id == "pi") {
            v = kPi;
            return true;
        }
        return error("unknown name '" + id + "' in an angle");
    }
};

struct Loader {
    Parser ps;
    std::vector<Reg> regs;
    std::vector<spz_op> ops;
    int nq = 0;

    const Reg *find(const std::string &name) const
    {
        for (const Reg &r : regs)
            if (r.name == name) return &r;
        return nullptr;
    }

    bool qubit(int &q)
    {
        std::string name;
        if (!ps.ident(name)) return false;
        const Reg *r = find(name);
        if (!r) return ps.error("unknown register '" + name + "'");
        if (!ps.peek('[')) return ps.error("a whole register as a gate argument is not supported");
        ps.expect('[');
        long i = 0;
        if (!ps.integer(i) || !ps.expect(']')) return false;
        if (i < 0 || i >= r->size) return ps.error("qubit index out of range for '" + name + "'");
        q = r->shift + (int)i;
        return true;
    }

    void add(int kind, int target, double th, double ph, double la, uint64_t ones)
    {
        spz_op o;
        memset(&o, 0, sizeof o);
        o.kind = kind;
        o.target = target;
        o.theta = th;
        o.phi = ph;
        o.lambda = la;
        o.ones = ones;
        // Controls::Single: classical control once the control is measured
        // (circuit.rs:570-578), as QuantumCircuit::cx/cp build it
        o.flags = ones ? SPZ_OPF_CLASSICAL : 0;
        o.u = 1.0;
        ops.push_back(o);
    }

    bool statement()
    {
        std::string word;
        if (!ps.ident(word)) return false;
        if (word == "OPENQASM") {
            ps.skip_ws();
            const char *s = ps.p;
            while (ps.p < ps.end && *ps.p != ';') ps.p++;
            std::string ver(s, ps.p);
            while (!ver.empty() && isspace((unsigned char)ver.back())) ver.pop_back();
            if (ver != "2.0") return ps.error("only OpenQASM 2.0 is supported");
            return ps.expect(';');
        }
        if (word == "include") {
            ps.skip_ws();
            const char *s = ps.p;
            while (ps.p < ps.end && *ps.p != ';' && *ps.p != '\n') ps.p++;
            std::string f(s, ps.p);
            if (f != "\"qelib1.inc\"") return ps.error("include of " + f + " is not supported");
            return ps.expect(';');
        }
        if (word == "qreg") {
            std::string name;
            long size = 0;
            if (!ps.ident(name) || !ps.expect('[') || !ps.integer(size) || !ps.expect(']') ||
                !ps.expect(';'))
                return false;
            if (find(name)) return ps.error("register '" + name + "' declared twice");
            if (size < 1) return ps.error("register size must be > 0");
            if (nq + size > 40) return ps.error("more than 40 qubits");
            regs.push_back({name, (int)size, nq});
            nq += (int)size;
            return true;
        }
        // gate application
        std::vector<double> args;
        if (ps.peek('(')) {
            ps.expect('(');
            if (!ps.peek(')')) {
                for (;;) {
                    double v;
                    if (!ps.expr(v)) return false;
                    args.push_back(v);
                    if (ps.peek(',')) { ps.expect(','); continue; }
                    break;
                }
            }
            if (!ps.expect(')')) return false;
        }
        std::vector<int> qs;
        for (;;) {
            int q;
            if (!qubit(q)) return false;
            qs.push_back(q);
            if (ps.peek(',')) { ps.expect(','); continue; }
            break;
        }
        if (!ps.expect(';')) return false;
        struct G { const char *name; int kind, nargs; };
        static const G single[] = {{"h", SPZ_H, 0}, {"x", SPZ_X, 0}, {"y", SPZ_Y, 0},
                                   {"z", SPZ_Z, 0}, {"rx", SPZ_RX, 1}, {"ry", SPZ_RY, 1},
                                   {"rz", SPZ_RZ, 1}, {"u", SPZ_U, 3}};
        for (const G &g : single) {
            if (word != g.name) continue;
            if ((int)args.size() != g.nargs)
                return ps.error("gate " + word + " takes " + std::to_string(g.nargs) + " angle(s)");
            for (int q : qs)
                add(g.kind, q, g.nargs > 0 ? args[0] : 0.0, g.nargs > 1 ? args[1] : 0.0,
                    g.nargs > 2 ? args[2] : 0.0, 0);
            return true;
        }
        if (word == "cx" || word == "cp") {
            const int na = word == "cp" ? 1 : 0;
            if ((int)args.size() != na)
                return ps.error("gate " + word + " takes " + std::to_string(na) + " angle(s)");
            if (qs.size() != 2) return ps.error("gate " + word + " takes a control and a target");
            if (qs[0] == qs[1]) return ps.error("control and target are the same qubit");
            add(word == "cp" ? SPZ_P : SPZ_X, qs[1], na ? args[0] : 0.0, 0.0, 0.0,
                1ull << qs[0]);
            return true;
        }
        return ps.error("gate '" + word + "' is not supported (h x y z rx ry rz u cp cx)");
    }

    bool run()
    {
        while (!ps.at_end())
            if (!statement()) return false;
        if (regs.empty()) return ps.error("no qreg declared");
        return true;
    }
};

}  // namespace

extern "C" int spz_qasm_parse(const char *src, uint64_t len, spz_op *ops, uint64_t cap,
                              uint64_t *n_ops, int *n_qubits)
{
    if (!src || !n_ops || !n_qubits) return spz::fail(SPZ_ERR_ARG, "null argument");
    if (cap && !ops) return spz::fail(SPZ_ERR_ARG, "null op array with a capacity");
    Loader ld;
    ld.ps.p = src;
    ld.ps.end = src + len;
    if (!ld.run()) return spz::fail(SPZ_ERR_ARG, "OpenQASM: " + ld.ps.err);
    *n_ops = ld.ops.size();
    *n_qubits = ld.nq;
    if (ld.ops.size() > cap) {
        if (cap) return spz::fail(SPZ_ERR_ARG, "op array too small");
        return SPZ_OK;  // size query
    }
    if (!ld.ops.empty()) memcpy(ops, ld.ops.data(), ld.ops.size() * sizeof(spz_op));
    return SPZ_OK;
}
