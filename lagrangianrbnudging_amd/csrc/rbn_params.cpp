// rbn_params.cpp — parameter handling of librbn: defaults, the "#define NAME value" reader,
// the Rayleigh-number map and validation.  Host code only.
//
// The reference configures everything through compile-time #define blocks
// (Trajectories/code/temppart_dump.c:15-31, Trajectories/code/mympi2.h:14-21,
//  Nudging/code/tempvel_nudge.c:13-31, Nudging/code/mympi2.h:15).  Those blocks are read
// here at run time: values are C constant expressions (e.g. `(-0.00004)`, `40*40`,
// `(1./3.)`), evaluated with C's int/double arithmetic rules, and may refer to earlier
// macros of the same file.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>

#include "rbn_internal.h"

namespace {

struct Val {
  bool ok = false;
  bool is_int = false;
  long long i = 0;
  double d = 0.0;
  double as_double() const { return is_int ? (double)i : d; }
};

// Recursive-descent evaluator for the constant expressions found in the #define blocks.
struct Expr {
  const char *p;
  const std::map<std::string, Val> *macros;
  int depth = 0;

  void ws() {
    while (*p && isspace((unsigned char)*p)) p++;
  }
  Val number() {
    Val v;
    const char *s = p;
    bool flt = false;
    while (isdigit((unsigned char)*p) || *p == '.' || *p == 'e' || *p == 'E' ||
           ((*p == '+' || *p == '-') && (p[-1] == 'e' || p[-1] == 'E'))) {
      if (*p == '.' || *p == 'e' || *p == 'E') flt = true;
      p++;
    }
    std::string tok(s, p);
    while (*p == 'f' || *p == 'F' || *p == 'l' || *p == 'L' || *p == 'u' || *p == 'U') {
      if (*p == 'f' || *p == 'F') flt = true;
      p++;
    }
    if (tok.empty()) return v;
    v.ok = true;
    if (flt) {
      v.d = strtod(tok.c_str(), nullptr);
    } else {
      v.is_int = true;
      v.i = strtoll(tok.c_str(), nullptr, 10);
    }
    return v;
  }
  Val primary() {
    ws();
    Val v;
    if (++depth > 64) return v;
    if (*p == '(') {
      p++;
      // C casts: (double) / (REAL) / (int) followed by an operand.
      const char *save = p;
      while (isspace((unsigned char)*p)) p++;
      const char *id0 = p;
      while (isalnum((unsigned char)*p) || *p == '_') p++;
      std::string word(id0, p);
      ws();
      if ((word == "double" || word == "REAL" || word == "float" || word == "int") && *p == ')') {
        p++;
        Val a = unary();
        if (!a.ok) return a;
        if (word == "int") {
          a.i = a.is_int ? a.i : (long long)a.d;
          a.is_int = true;
        } else {
          a.d = a.as_double();
          a.is_int = false;
        }
        depth--;
        return a;
      }
      p = save;
      v = expr();
      ws();
      if (*p != ')') return Val{};
      p++;
    } else if (isdigit((unsigned char)*p) || *p == '.') {
      v = number();
    } else if (isalpha((unsigned char)*p) || *p == '_') {
      const char *s = p;
      while (isalnum((unsigned char)*p) || *p == '_') p++;
      auto it = macros->find(std::string(s, p));
      if (it != macros->end()) v = it->second;
    }
    depth--;
    return v;
  }
  Val unary() {
    ws();
    if (*p == '-') {
      p++;
      Val a = unary();
      if (a.is_int) a.i = -a.i; else a.d = -a.d;
      return a;
    }
    if (*p == '+') {
      p++;
      return unary();
    }
    return primary();
  }
  static Val arith(const Val &a, const Val &b, char op) {
    Val r;
    r.ok = a.ok && b.ok;
    if (!r.ok) return r;
    if (a.is_int && b.is_int) {
      r.is_int = true;
      switch (op) {
        case '*': r.i = a.i * b.i; break;
        case '/': if (b.i == 0) { r.ok = false; return r; } r.i = a.i / b.i; break;
        case '+': r.i = a.i + b.i; break;
        default: r.i = a.i - b.i; break;
      }
    } else {
      double x = a.as_double(), y = b.as_double();
      switch (op) {
        case '*': r.d = x * y; break;
        case '/': r.d = x / y; break;
        case '+': r.d = x + y; break;
        default: r.d = x - y; break;
      }
    }
    return r;
  }
  Val term() {
    Val a = unary();
    for (;;) {
      ws();
      if (*p == '*' || *p == '/') {
        char op = *p++;
        a = arith(a, unary(), op);
      } else {
        return a;
      }
    }
  }
  Val expr() {
    Val a = term();
    for (;;) {
      ws();
      if (*p == '+' || *p == '-') {
        char op = *p++;
        a = arith(a, term(), op);
      } else {
        return a;
      }
    }
  }
};

thread_local std::string g_err;

int set_err(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

struct Field {
  const char *name;
  enum { I, D } kind;
  size_t off;
};

#define FI(n, m) {n, Field::I, offsetof(rbn_params, m)}
#define FD(n, m) {n, Field::D, offsetof(rbn_params, m)}
const Field kFields[] = {
    FI("nx", nx),
    FI("ny", ny),
    FI("scratch", scratch),
    FI("contin", contin),
    FI("nstart", nstart),
    FI("nsteps", nsteps),
    FI("nout", nout),
    FI("noutconfig", noutconfig),
    FI("ntempread", ntempread),
    FD("rhom1", rhom1),
    FD("relax", relax),
    FD("relaxg", relaxg),
    FD("Tup", Tup),
    FD("Tdown", Tdown),
    FD("gravity", gravity),
    FD("alpha", alpha),
    FI("Particle", particle),
    FI("Particle_scratch", particle_scratch),
    FI("Nprtcl", nprtcl),
    FI("tracer", tracer),
    FI("heavyprtcl", heavyprtcl),
    FD("taulag", taulag),
    FD("beta", beta),
    FI("Prtcl_Writefull", prtcl_writefull),
    FI("rbn_variant", variant),
    FI("rbn_precision", precision),
    FI("rbn_deposit", deposit),
    FI("rbn_fix_q3", fix_q3),
    FI("rbn_ndiag", ndiag),
    FD("rbn_T0_shift", T0_shift),
    FD("rbn_rho0_shift", rho0_shift),
    FD("rbn_heavy_init_factor", heavy_init_factor),
};
#undef FI
#undef FD

const Field *find_field(const char *name) {
  for (const Field &f : kFields)
    if (strcmp(f.name, name) == 0) return &f;
  return nullptr;
}

void store(rbn_params *p, const Field *f, double value) {
  char *base = reinterpret_cast<char *>(p);
  if (f->kind == Field::I)
    *reinterpret_cast<int *>(base + f->off) = (int)value;
  else
    *reinterpret_cast<double *>(base + f->off) = value;
}

}  // namespace

namespace rbn {
const char *thread_error() { return g_err.c_str(); }
int thread_fail(int code, const char *msg) { return set_err(code, msg); }
}  // namespace rbn

extern "C" {

void rbn_params_default(rbn_params *p) {
  if (!p) return;
  memset(p, 0, sizeof *p);
  // Trajectories/code/temppart_dump.c:15-31
  p->nx = 1680;
  p->ny = 840;
  p->scratch = 0;
  p->contin = 0;
  p->nstart = 0;
  p->nsteps = 5000000;
  p->nout = 20000;
  p->noutconfig = 250;
  p->ntempread = 250;  // Nudging/code/tempvel_nudge.c:23
  p->rhom1 = 1.0;
  p->relax = 0.502;
  p->relaxg = 0.502;
  p->Tup = -0.01;
  p->Tdown = 0.01;
  p->gravity = -0.00004;
  p->alpha = 0.01;  // Nudging/code/tempvel_nudge.c:13
  p->particle = 1;
  p->particle_scratch = 1;
  // Trajectories/code/mympi2.h:14-21
  p->nprtcl = 1600;
  p->tracer = 0;
  p->heavyprtcl = 1;
  p->taulag = 2150.0;
  p->beta = 0.1;
  p->prtcl_writefull = 1;
  // options
  p->variant = RBN_VARIANT_TRAJECTORIES;
  p->precision = RBN_FP64;
  p->deposit = RBN_DEPOSIT_BOX7;
  p->fix_q3 = 0;
  p->ndiag = 0;
  p->T0_shift = 0.0;
  p->rho0_shift = 1.0;
  p->heavy_init_factor = 0.98;  // Trajectories/code/mympi2.h:205-207
}

int rbn_params_set(rbn_params *p, const char *name, double value) {
  if (!p || !name) return set_err(RBN_EINVAL, "rbn_params_set: null argument");
  const Field *f = find_field(name);
  if (!f) return set_err(RBN_EINVAL, std::string("unknown parameter ") + name);
  store(p, f, value);
  return RBN_OK;
}

int rbn_params_parse_text(const char *text, rbn_params *p) {
  if (!text || !p) return set_err(RBN_EINVAL, "rbn_params_parse_text: null argument");
  std::map<std::string, Val> macros;
  bool saw_nudging = false, saw_variant = false;
  std::string src(text);
  // Strip /* */ comments (they may span lines); // comments are removed per line.
  for (size_t a = src.find("/*"); a != std::string::npos; a = src.find("/*", a)) {
    size_t b = src.find("*/", a + 2);
    if (b == std::string::npos) {
      src.erase(a);
      break;
    }
    for (size_t k = a; k < b + 2; k++)
      if (src[k] != '\n') src[k] = ' ';
  }
  size_t pos = 0;
  int lineno = 0;
  while (pos < src.size()) {
    size_t eol = src.find('\n', pos);
    if (eol == std::string::npos) eol = src.size();
    std::string line = src.substr(pos, eol - pos);
    pos = eol + 1;
    lineno++;
    size_t c = line.find("//");
    if (c != std::string::npos) line.erase(c);
    const char *s = line.c_str();
    while (isspace((unsigned char)*s)) s++;
    if (*s != '#') continue;
    s++;
    while (isspace((unsigned char)*s)) s++;
    if (strncmp(s, "define", 6) != 0 || !isspace((unsigned char)s[6])) continue;
    s += 6;
    while (isspace((unsigned char)*s)) s++;
    const char *n0 = s;
    while (isalnum((unsigned char)*s) || *s == '_') s++;
    std::string name(n0, s);
    if (name.empty() || *s == '(') continue;  // function-like macros are not parameters
    Expr e{s, &macros};
    Val v = e.expr();
    e.ws();
    if (!v.ok || *e.p != '\0') {
      // Not a constant expression (e.g. `(!me)`, `sqrt((cs2))`); only an error when the
      // name is a parameter.
      if (find_field(name.c_str())) {
        char msg[256];
        snprintf(msg, sizeof msg, "line %d: cannot evaluate #define %s", lineno, name.c_str());
        return set_err(RBN_EPARSE, msg);
      }
      continue;
    }
    macros[name] = v;
    const Field *f = find_field(name.c_str());
    if (f) store(p, f, v.as_double());
    if (name == "ntempread" || name == "alpha") saw_nudging = true;
    if (name == "rbn_variant") saw_variant = true;
  }
  if (saw_nudging && !saw_variant) p->variant = RBN_VARIANT_NUDGING;
  return RBN_OK;
}

int rbn_params_parse(const char *path, rbn_params *p) {
  if (!path || !p) return set_err(RBN_EINVAL, "rbn_params_parse: null argument");
  FILE *fp = fopen(path, "rb");
  if (!fp) return set_err(RBN_EIO, std::string("cannot open ") + path);
  std::string text;
  char buf[65536];
  size_t n;
  while ((n = fread(buf, 1, sizeof buf, fp)) > 0) text.append(buf, n);
  fclose(fp);
  return rbn_params_parse_text(text.c_str(), p);
}

// media() (Nudging/code/tempvel_nudge.c:1262-1283).
double rbn_params_rayleigh(const rbn_params *p) {
  const double cs2 = (1. / 3.);
  double nu = cs2 * (p->relax - 0.5);
  double kappa = cs2 * (p->relaxg - 0.5);
  double Delta = (p->Tdown - p->Tup);
  double L = (double)(p->ny - 1);
  return -(p->gravity / p->rhom1) * L * L * L * Delta / (nu * kappa);
}

int rbn_params_from_ra(rbn_params *p, double Ra, double Pr, double u_ff) {
  if (!p || !(Ra > 0) || !(Pr > 0) || !(u_ff > 0) || p->ny < 2)
    return set_err(RBN_EINVAL, "rbn_params_from_ra: need Ra, Pr, u_ff > 0 and ny >= 2");
  double dT = p->Tdown - p->Tup;
  if (!(dT != 0.0)) return set_err(RBN_EINVAL, "rbn_params_from_ra: Tdown == Tup");
  double L = (double)(p->ny - 1);
  // u_ff^2 = |gravity/rhom1| dT L ; Ra = u_ff^2 L^2 / (nu kappa) ; Pr = nu / kappa.
  double g_over_rho = u_ff * u_ff / (fabs(dT) * L);
  double nu = u_ff * L * sqrt(Pr / Ra);
  double kappa = nu / Pr;
  p->relax = 0.5 + 3.0 * nu;
  p->relaxg = 0.5 + 3.0 * kappa;
  p->gravity = -(dT > 0 ? 1.0 : -1.0) * g_over_rho * p->rhom1;
  return RBN_OK;
}

int rbn_params_check(const rbn_params *p) {
  char msg[256];
  if (!p) return set_err(RBN_EINVAL, "null params");
  if (p->nx < 8 || p->ny < 4) return set_err(RBN_EINVAL, "lattice must be at least 8 x 4");
  if ((long long)p->nx * (long long)p->ny > (1LL << 31))
    return set_err(RBN_ERANGE, "lattice larger than 2^31 nodes");
  if (!(p->relax > 0.5) || !(p->relaxg > 0.5))
    return set_err(RBN_EINVAL, "relax and relaxg must exceed 0.5");
  if (p->precision != RBN_FP64 && p->precision != RBN_FP32)
    return set_err(RBN_EINVAL, "precision must be RBN_FP64 or RBN_FP32");
  if (p->variant != RBN_VARIANT_TRAJECTORIES && p->variant != RBN_VARIANT_NUDGING)
    return set_err(RBN_EINVAL, "unknown variant");
  if (p->noutconfig < 1) return set_err(RBN_EINVAL, "noutconfig must be >= 1");
  if (p->deposit != RBN_DEPOSIT_BOX7 && p->deposit != RBN_DEPOSIT_BILINEAR)
    return set_err(RBN_EINVAL, "deposit must be RBN_DEPOSIT_BOX7 or RBN_DEPOSIT_BILINEAR");
  if (p->variant == RBN_VARIANT_NUDGING) {
    if (p->ntempread < 2) return set_err(RBN_EINVAL, "ntempread must be >= 2");
    if (p->contin != 0) return set_err(RBN_EINVAL, "contin = 1 (restart of a nudged run) is not supported");
    // Record istep+ntempread-1 is read on steps istep % ntempread == 1 (N:1390,1415), so
    // the run must start on an interval boundary.
    if (p->nstart % p->ntempread != 0) {
      snprintf(msg, sizeof msg, "nstart (%d) must be a multiple of ntempread (%d)", p->nstart, p->ntempread);
      return set_err(RBN_EINVAL, msg);
    }
  }
  if (p->precision == RBN_FP32 && !(std::isfinite(p->T0_shift) && std::isfinite(p->rho0_shift)))
    return set_err(RBN_EINVAL, "FP32 shifts must be finite");
  return RBN_OK;
}

}  // extern "C"
