// kppgen — mechanism compiler for the MI355X batched Rosenbrock solver.
//
// Reads a mechanism written in the KPP language (.def/.spc/.eqn/.kpp files)
// and emits, from the same expression trees:
//   <root>_model.h  plain C model functions with KPP's C ABI semantics
//                   (Fun, Jac_SP, KppSolve, Update_RCONST, LU pattern arrays);
//                   consumed by the CPU oracle.
//   <root>_hip.h    HIP __device__ versions of Fun, Jac_SP (fused with the
//                   Ghimj build), an unrolled KppDecomp, KppSolve and the rate
//                   constants, templated on an element accessor so the same
//                   code runs on SoA global memory, LDS or register arrays.
//   <root>_stats.json  sizes and exact operation counts of every routine.
//
// What has to match KPP (and why) — each point restates the cited reference:
//  * species numbering: first occurrence in the equations
//    (src/scanner.c:712-766 ProcessTerm), declaration order as the "code";
//  * variable ordering: reactive first, then by code (src/kpp.c:73-88
//    CodeRCmp, ReorderSpecies UNSORT :273-402), then Markowitz
//    (src/kpp.c:220-271 BestSparsity, first strict minimum of (cnz-1)(lnz-1));
//  * Jacobian and LU patterns (src/kpp.c:113-158, src/gen.c:331-356 NonZero);
//  * stoichiometric coefficients stored as float and printed with "%g"
//    (src/scanner.c:712-766, src/code_c.c:66), NUMBER rates folded in as
//    "%g"-of-float constants (src/gen.c:118-125 RConst);
//  * the simplification rules of the expression builder (src/code.c:519-597)
//    and left-to-right association of the A[], Vdot[], B[], JVS[] and solve
//    sums (src/gen.c:713-761, 1239-1284, 1991-2024), because the association
//    order fixes the floating-point result.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <map>
#include <set>
#include <array>
#include <sstream>
#include <string>
#include <vector>

namespace {

[[noreturn]] void die(const std::string& m) {
  fprintf(stderr, "kppgen: error: %s\n", m.c_str());
  exit(2);
}

std::string upper(std::string s) {
  for (auto& c : s) c = (char)toupper((unsigned char)c);
  return s;
}
bool eq_nocase(const std::string& a, const std::string& b) { return upper(a) == upper(b); }
std::string trim(const std::string& s) {
  size_t b = 0, e = s.size();
  while (b < e && isspace((unsigned char)s[b])) b++;
  while (e > b && isspace((unsigned char)s[e - 1])) e--;
  return s.substr(b, e - b);
}
bool file_exists(const std::string& p) {
  std::ifstream f(p);
  return f.good();
}
std::string read_file(const std::string& p) {
  std::ifstream f(p, std::ios::binary);
  if (!f) die("cannot open " + p);
  std::stringstream ss;
  ss << f.rdbuf();
  return ss.str();
}
std::string dirname_of(const std::string& p) {
  auto k = p.find_last_of('/');
  return k == std::string::npos ? std::string(".") : p.substr(0, k);
}

// ---------------------------------------------------------------------------
// Lexical pass: comments, #INCLUDE, #INLINE, and splitting into command
// sections.  Mirrors the scanner states of src/scan.l:112-296 for the subset
// of commands the solver needs: '{...}' and '//' comments are dropped outside
// inline code, a '#WORD' starts a command, inline code is kept verbatim up to
// #ENDINLINE.
// ---------------------------------------------------------------------------
struct Section {
  std::string cmd;   // upper-case command name, without '#'
  std::string body;  // text up to the next command
};

struct Scan {
  std::vector<std::string> search;  // include search directories
  std::vector<Section> sections;
  std::map<std::string, std::string> inlines;  // inline context -> code
  int depth = 0;

  std::string resolve(const std::string& name, const std::string& from_dir) {
    std::vector<std::string> cands;
    auto add = [&](const std::string& d) {
      cands.push_back(d + "/" + name);
      // KPP appends its default extension when the name has none
      // (src/scanutil.c:142-200 FileName).
      if (name.find('.') == std::string::npos) cands.push_back(d + "/" + name + ".kpp");
    };
    if (!name.empty() && name[0] == '/') {
      cands.push_back(name);
    } else {
      add(from_dir);
      for (auto& d : search) add(d);
    }
    for (auto& c : cands)
      if (file_exists(c)) return c;
    die("cannot resolve include '" + name + "' from " + from_dir);
  }

  static bool is_tok_char(char c) {  // STRING of src/scan.l:101: [^ \t\n{}#;]
    return !(c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '{' || c == '}' || c == '#' ||
             c == ';');
  }

  void file(const std::string& path) {
    if (++depth > 16) die("include nesting too deep at " + path);
    std::string t = read_file(path);
    std::string dir = dirname_of(path);
    size_t i = 0, n = t.size();
    auto read_token = [&](size_t& k) {
      while (k < n && (t[k] == ' ' || t[k] == '\t')) k++;
      size_t b = k;
      while (k < n && is_tok_char(t[k])) k++;
      return t.substr(b, k - b);
    };
    if (sections.empty()) sections.push_back({"", ""});
    while (i < n) {
      char c = t[i];
      if (c == '{') {
        size_t j = t.find('}', i + 1);
        i = (j == std::string::npos) ? n : j + 1;
        continue;
      }
      if (c == '/' && i + 1 < n && t[i + 1] == '/') {
        size_t j = t.find('\n', i);
        i = (j == std::string::npos) ? n : j;
        continue;
      }
      if (c == '#') {
        size_t k = i + 1;
        std::string w = upper(read_token(k));
        if (w == "INLINE") {
          std::string ctx = upper(read_token(k));
          // code starts after the context name; ends at "#ENDINLINE"
          size_t b = k, e = b;
          for (;;) {
            e = t.find('#', e);
            if (e == std::string::npos) die("unterminated #INLINE " + ctx + " in " + path);
            size_t q = e + 1;
            std::string w2 = upper(read_token(q));
            if (w2 == "ENDINLINE") {
              inlines[ctx] += t.substr(b, e - b);
              k = q;
              break;
            }
            e = e + 1;
          }
          i = k;
          continue;
        }
        if (w == "INCLUDE" || w == "MODEL") {
          std::string f = read_token(k);
          std::string p = (w == "MODEL") ? resolve(f + ".def", dir) : resolve(f, dir);
          i = k;
          file(p);
          continue;
        }
        sections.push_back({w, ""});
        i = k;
        continue;
      }
      sections.back().body += c;
      i++;
    }
    depth--;
  }
};

// ---------------------------------------------------------------------------
// Mechanism tables (src/gdata.h): species codes, stoichiometry, rates.
// ---------------------------------------------------------------------------
enum SpcType { VAR_SPC = 1, FIX_SPC = 2, DUMMY_SPC = 3 };
enum RateType { NUMBER = 1, EXPRESSION = 2, PHOTO = 3 };

struct SpeciesDef {
  std::string name;
  int type;
  std::string ival;  // initial value text from #INITVALUES
};

struct Rate {
  std::string label;
  int type;
  double f = 0;      // NUMBER value
  std::string expr;  // EXPRESSION / PHOTO text (tokens concatenated)
};

constexpr int CODE_HV = 1000000 - 1;   // stands for MAX_SPECIES-1 (KPP's "HV")
constexpr int CODE_PROD = 1000000 - 2;  // MAX_SPECIES-2 ("PROD")

struct Mech {
  std::string root;
  std::vector<SpeciesDef> table;  // index = code; [0] is KPP's "???" dummy
  SpeciesDef hv{"HV", DUMMY_SPC, ""}, prod{"PROD", DUMMY_SPC, ""};
  double cfactor = 1.0;
  std::string varDefault = "0", fixDefault = "0";
  bool cfactor_set = false;

  // per appearance-order species index ("crtSpec"): code, reactive flag
  std::vector<int> code;
  std::vector<int> reactive;
  std::map<int, int> reverse;  // code -> appearance index
  // stoichiometry [species appearance index][equation], float like KPP
  std::vector<std::vector<float>> sl, sr, st;
  std::vector<Rate> rates;
  std::vector<std::string> families;  // #FAMILIES names, in declaration order

  SpeciesDef& spc(int c) {
    if (c == CODE_HV) return hv;
    if (c == CODE_PROD) return prod;
    return table[c];
  }
  int find_species(const std::string& n) {
    for (size_t i = 0; i < table.size(); i++)
      if (eq_nocase(table[i].name, n)) return (int)i;
    if (eq_nocase(n, "HV")) return CODE_HV;
    if (eq_nocase(n, "PROD")) return CODE_PROD;
    return -1;
  }
};

std::vector<std::string> split_semis(const std::string& body) {
  std::vector<std::string> out;
  std::string cur;
  for (char c : body) {
    if (c == ';') {
      out.push_back(cur);
      cur.clear();
    } else {
      cur += c;
    }
  }
  if (!trim(cur).empty()) out.push_back(cur);
  return out;
}

std::vector<std::string> split_plus(const std::string& body) {
  std::vector<std::string> out(1);
  for (char c : body) {
    if (c == '+') out.emplace_back();
    else out.back() += c;
  }
  return out;
}

void declare_species(Mech& m, const std::string& body, int type) {
  for (auto& item : split_semis(body)) {
    std::string s = trim(item);
    if (s.empty()) continue;
    auto eqp = s.find('=');
    std::string name = trim(eqp == std::string::npos ? s : s.substr(0, eqp));
    if (name.empty()) continue;
    if (m.find_species(name) >= 0) die("multiple declaration for species " + name);
    m.table.push_back({name, type, ""});
  }
}

void set_type(Mech& m, const std::string& body, int type) {  // #SETVAR / #SETFIX
  for (auto& item : split_semis(body)) {
    std::string n = trim(item);
    if (n.empty()) continue;
    if (eq_nocase(n, "VAR_SPEC") || eq_nocase(n, "FIX_SPEC") || eq_nocase(n, "ALL_SPEC")) {
      int from = eq_nocase(n, "VAR_SPEC") ? VAR_SPC : eq_nocase(n, "FIX_SPEC") ? FIX_SPC : 0;
      for (size_t i = 1; i < m.table.size(); i++)
        if (!from || m.table[i].type == from) m.table[i].type = type;
      continue;
    }
    int c = m.find_species(n);
    if (c < 0) die("undefined species " + n);
    m.spc(c).type = type;
  }
}

void init_values(Mech& m, const std::string& body) {
  for (auto& item : split_semis(body)) {
    std::string s = trim(item);
    if (s.empty()) continue;
    auto eqp = s.find('=');
    if (eqp == std::string::npos) die("bad #INITVALUES entry: " + s);
    std::string n = trim(s.substr(0, eqp)), v = trim(s.substr(eqp + 1));
    if (eq_nocase(n, "CFACTOR")) {
      m.cfactor = strtod(v.c_str(), nullptr);
      m.cfactor_set = true;
    } else if (eq_nocase(n, "VAR_SPEC")) {
      m.varDefault = v;
    } else if (eq_nocase(n, "FIX_SPEC")) {
      m.fixDefault = v;
    } else if (eq_nocase(n, "ALL_SPEC")) {
      m.varDefault = m.fixDefault = v;
    } else {
      int c = m.find_species(n);
      if (c < 0) die("undefined species in #INITVALUES: " + n);
      m.spc(c).ival = v;
    }
  }
}

// One equation "<tag> lhs = rhs : rate" (src/scan.y:307-359).
void parse_equation(Mech& m, const std::string& text) {
  std::string s = trim(text);
  if (s.empty()) return;
  std::string label = "          ";
  size_t i = 0;
  if (s[0] == '<') {
    auto k = s.find('>');
    if (k == std::string::npos) die("bad equation tag: " + s);
    label = trim(s.substr(1, k - 1));
    i = k + 1;
  }
  auto eqp = s.find('=', i);
  auto colon = s.find(':', i);
  if (eqp == std::string::npos || colon == std::string::npos || colon < eqp)
    die("bad equation: " + s);
  std::string lhs = s.substr(i, eqp - i), rhs = s.substr(eqp + 1, colon - eqp - 1);
  std::string rate_raw = s.substr(colon + 1);
  int eqn = (int)m.rates.size();
  for (auto& row : m.sl) row.push_back(0.f);
  for (auto& row : m.sr) row.push_back(0.f);
  for (auto& row : m.st) row.push_back(0.f);
  bool is_photo = false;

  auto process = [&](int side, const std::string& sign, const std::string& coef,
                     const std::string& name) {
    int c = m.find_species(name);
    if (c < 0) die("undefined species '" + name + "' in equation <" + label + ">");
    if (eq_nocase(name, "HV")) is_photo = true;
    int k;
    auto it = m.reverse.find(c);
    if (it == m.reverse.end()) {
      k = (int)m.code.size();
      m.code.push_back(c);
      m.reactive.push_back(0);
      m.reverse[c] = k;
      m.sl.emplace_back(eqn + 1, 0.f);
      m.sr.emplace_back(eqn + 1, 0.f);
      m.st.emplace_back(eqn + 1, 0.f);
    } else {
      k = it->second;
    }
    double val = strtod((sign + coef).c_str(), nullptr);
    if (side == 0) {
      m.sl[k][eqn] += val;
      m.st[k][eqn] -= val;
      m.reactive[k] = 1;
    } else {
      m.sr[k][eqn] += val;
      m.st[k][eqn] += val;
    }
  };
  auto terms = [&](const std::string& e, int side) {
    size_t p = 0, n = e.size();
    std::string sign = "+";
    while (p < n) {
      char c = e[p];
      if (isspace((unsigned char)c)) { p++; continue; }
      if (c == '+' || c == '-') { sign = std::string(1, c); p++; continue; }
      std::string coef = "1";
      if (isdigit((unsigned char)c) || c == '.') {  // UFLOAT of src/scan.l:112
        size_t b = p;
        while (p < n && (isdigit((unsigned char)e[p]) || e[p] == '.')) p++;
        if (p < n && (e[p] == 'e' || e[p] == 'E' || e[p] == 'd' || e[p] == 'D') && p + 1 < n &&
            (isdigit((unsigned char)e[p + 1]) ||
             ((e[p + 1] == '+' || e[p + 1] == '-') && p + 2 < n && isdigit((unsigned char)e[p + 2])))) {
          p += 2;
          while (p < n && isdigit((unsigned char)e[p])) p++;
        }
        coef = e.substr(b, p - b);
        while (p < n && isspace((unsigned char)e[p])) p++;
      }
      if (p < n && (isalpha((unsigned char)e[p]) || e[p] == '_')) {
        size_t b = p;
        while (p < n && (isalnum((unsigned char)e[p]) || e[p] == '_')) p++;
        process(side, sign, coef, e.substr(b, p - b));
        sign = "+";
        continue;
      }
      die("cannot parse equation term in <" + label + ">: '" + e + "'");
    }
  };
  terms(lhs, 0);
  terms(rhs, 1);
  // Rate text: whitespace-separated tokens concatenated (src/scan.y:324-330).
  std::string rate;
  for (char c : rate_raw)
    if (!isspace((unsigned char)c)) rate += c;
  Rate r;
  r.label = label;
  if (is_photo) {
    r.type = PHOTO;
    r.expr = rate;
  } else {
    // NUMBER iff the whole text is one floating literal (src/scanner.c:604-628)
    char* end = nullptr;
    double f = strtod(rate.c_str(), &end);
    if (end != rate.c_str() && *end == 0) {
      r.type = NUMBER;
      r.f = f;
    } else {
      r.type = EXPRESSION;
      r.expr = rate;
    }
  }
  m.rates.push_back(r);
}

// #FAMILIES: production/loss families (src/scan.y:244-268, src/scanner.c:
// 921-1097).  "Pname : m1 + 2 m2;" or "Lname : ...": for every equation
// parsed so far, the family's loss L = sum coeff*Stoich_Left and production
// P = sum coeff*Stoich_Right over its members (floats, as the reference's
// Loss_Coeff/Prod_Coeff); where P - L (production family) or L - P (loss
// family) is positive, a variable species named after the family (declared
// on first use) is made a product of the equation with that coefficient,
// printed with "%f" and read back as the reference does.  A member's
// coefficient is truncated to an integer (the reference stores
// (unsigned char)atof(nr)).  Families declared before the equations see no
// equations, as in the reference.
void declare_families(Mech& m, const std::string& body) {
  for (auto& item : split_semis(body)) {
    std::string s = trim(item);
    if (s.empty()) continue;
    auto colon = s.find(':');
    if (colon == std::string::npos) die("bad #FAMILIES entry: " + s);
    std::string fam = trim(s.substr(0, colon));
    for (auto& f : m.families)
      if (eq_nocase(f, fam)) die("multiple declaration for family " + fam);
    int type;
    if (!fam.empty() && fam[0] == 'P') type = 1;       // PROD_FAM
    else if (!fam.empty() && fam[0] == 'L') type = 2;  // LOSS_FAM
    else die(fam + " must be declared as a P/L term with appropriate prefix (P,L)");
    const int neq = (int)m.rates.size();
    std::vector<float> loss(neq, 0.f), prod(neq, 0.f);
    std::string rest = s.substr(colon + 1);
    for (auto& term : split_plus(rest)) {
      std::string t = trim(term), nr = "1", name = t;
      size_t p = 0;
      while (p < t.size() && (isdigit((unsigned char)t[p]) || t[p] == '.')) p++;
      if (p > 0) {
        nr = t.substr(0, p);
        name = trim(t.substr(p));
      }
      const int c = m.find_species(name);
      if (c < 0) die("undefined member " + name + " of family " + fam);
      const float coeff = (float)(unsigned char)atof(nr.c_str());
      auto it = m.reverse.find(c);
      if (it == m.reverse.end()) continue;  // in no equation
      const int k = it->second;
      for (int i = 0; i < neq; i++) {
        if (m.sl[k][i] > 0) loss[i] += m.sl[k][i] * coeff;
        if (m.sr[k][i] > 0) prod[i] += m.sr[k][i] * coeff;
      }
    }
    m.families.push_back(fam);
    for (int i = 0; i < neq; i++) {
      const float net = (type == 1) ? prod[i] - loss[i] : loss[i] - prod[i];
      if (!(net > 0.f)) continue;
      int c = m.find_species(fam);
      if (c < 0) {
        m.table.push_back({fam, VAR_SPC, ""});
        c = (int)m.table.size() - 1;
      }
      int k;
      auto it = m.reverse.find(c);
      if (it == m.reverse.end()) {
        k = (int)m.code.size();
        m.code.push_back(c);
        m.reactive.push_back(0);
        m.reverse[c] = k;
        m.sl.emplace_back(neq, 0.f);
        m.sr.emplace_back(neq, 0.f);
        m.st.emplace_back(neq, 0.f);
      } else {
        k = it->second;
      }
      char buf[64];
      snprintf(buf, sizeof buf, "+%f", (double)net);
      const double val = strtod(buf, nullptr);
      m.sr[k][i] = (float)val;
      m.st[k][i] = (float)val;
    }
  }
}

// ---------------------------------------------------------------------------
// Structural analysis (src/kpp.c).  After reorder(), species index k has
// code[k], stoichiometry sl/sr/st[k][*]; [0,nvar) are variable, then fixed.
// ---------------------------------------------------------------------------
struct Structure {
  int nvar = 0, nfix = 0, nspc = 0, nreact = 0;
  std::vector<std::vector<int>> J, LU;  // nvar x nvar
  std::vector<int> lu_icol, lu_irow, lu_crow, lu_diag;
  int nonzero = 0, lu_nonzero = 0;
};

void struct_j(const Mech& m, int nvar, std::vector<std::vector<int>>& J) {
  J.assign(nvar, std::vector<int>(nvar, 0));
  int ne = (int)m.rates.size();
  for (int i = 0; i < nvar; i++) J[i][i] = 1;
  for (int i = 0; i < nvar; i++)
    for (int j = 0; j < nvar; j++)
      for (int k = 0; k < ne; k++)
        if (m.st[i][k] * (m.sl[j][k] != 0 ? 1 : 0) != 0.0f) {
          J[i][j] = 1;
          break;
        }
}

void lu_fill(std::vector<std::vector<int>>& L, int n) {  // src/kpp.c:133-158
  for (int j = 0; j < n - 1; j++)
    for (int i = j + 1; i < n; i++)
      if (L[i][j])
        for (int k = j; k < n; k++)
          if (L[j][k] != 0) L[i][k] = 1;
}

// Rebuild the per-index tables in a new species order given as a list of
// appearance indices (variables first, then fixed, then dummies).
void permute(Mech& m, const std::vector<int>& order) {
  auto P = [&](auto& v) {
    auto old = v;
    for (size_t k = 0; k < order.size(); k++) v[k] = old[order[k]];
  };
  P(m.code);
  P(m.reactive);
  P(m.sl);
  P(m.sr);
  P(m.st);
  m.reverse.clear();
  for (size_t k = 0; k < m.code.size(); k++) m.reverse[m.code[k]] = (int)k;
}

Structure reorder(Mech& m, bool use_reorder) {
  Structure S;
  int n = (int)m.code.size();
  std::vector<int> var, fix, dummy;
  for (int k = 0; k < n; k++) {
    int t = m.spc(m.code[k]).type;
    if (t == VAR_SPC) var.push_back(k);
    else if (t == FIX_SPC) fix.push_back(k);
    else dummy.push_back(k);
  }
  // UNSORT: reactive variables first, each group by code (src/kpp.c:73-88);
  // fixed species by code.
  std::sort(var.begin(), var.end(), [&](int a, int b) {
    if (m.reactive[a] != m.reactive[b]) return m.reactive[a] > m.reactive[b];
    return m.code[a] < m.code[b];
  });
  std::sort(fix.begin(), fix.end(), [&](int a, int b) { return m.code[a] < m.code[b]; });
  std::vector<int> order;
  order.insert(order.end(), var.begin(), var.end());
  order.insert(order.end(), fix.begin(), fix.end());
  order.insert(order.end(), dummy.begin(), dummy.end());
  permute(m, order);
  S.nvar = (int)var.size();
  S.nfix = (int)fix.size();
  S.nspc = S.nvar + S.nfix;
  S.nreact = (int)m.rates.size();

  if (use_reorder && S.nvar > 0) {
    // BestSparsity (src/kpp.c:220-271): greedy Markowitz on the symbolic
    // structure with explicit fill, first strict minimum wins.
    std::vector<std::vector<int>> L;
    struct_j(m, S.nvar, L);
    int nv = S.nvar;
    std::vector<int> best(nv);
    for (int i = 0; i < nv; i++) best[i] = i;
    for (int s = 0; s < nv - 1; s++) {
      long bestv = 1L << 62;
      int bi = 0;
      for (int i = s; i < nv; i++) {
        long cnz = 0, lnz = 0;
        for (int j = s; j < nv; j++) {
          cnz += L[i][j] ? 1 : 0;
          lnz += L[j][i] ? 1 : 0;
        }
        long crt = (cnz - 1) * (lnz - 1);
        if (crt < bestv) {
          bestv = crt;
          bi = i;
        }
      }
      std::swap(L[s], L[bi]);
      for (int i = 0; i < nv; i++) std::swap(L[i][s], L[i][bi]);
      std::swap(best[s], best[bi]);
      for (int i = s + 1; i < nv; i++)
        if (L[i][s])
          for (int k = s; k < nv; k++) L[i][k] += L[s][k];
    }
    std::vector<int> ord2;
    for (int i = 0; i < nv; i++) ord2.push_back(best[i]);
    for (int k = nv; k < n; k++) ord2.push_back(k);
    permute(m, ord2);
  }
  struct_j(m, S.nvar, S.J);
  S.LU = S.J;
  lu_fill(S.LU, S.nvar);
  // NonZero (src/gen.c:331-356), row-major, diagonal always present.
  auto nonzero = [&](const std::vector<std::vector<int>>& M, std::vector<int>* irow,
                     std::vector<int>* icol, std::vector<int>* crow, std::vector<int>* diag) {
    int ne = 0;
    for (int i = 0; i < S.nvar; i++) {
      if (crow) crow->push_back(ne);
      for (int j = 0; j < S.nvar; j++) {
        if (i == j || M[i][j]) {
          if (irow) irow->push_back(i);
          if (icol) icol->push_back(j);
          ne++;
        }
        if (i == j && diag) diag->push_back(ne - 1);
      }
    }
    if (crow) crow->push_back(ne);
    if (diag) diag->push_back(ne);
    return ne;
  };
  S.nonzero = nonzero(S.J, nullptr, nullptr, nullptr, nullptr);
  S.lu_nonzero = nonzero(S.LU, &S.lu_irow, &S.lu_icol, &S.lu_crow, &S.lu_diag);
  return S;
}

// ---------------------------------------------------------------------------
// Expression nodes with the simplification rules of src/code.c:519-597.
// A node is a constant (|value| as float plus a sign), an element reference,
// or a binary operation.  Printing hoists leaf signs: in IEEE round-to-nearest
// a + (-b) == a - b and (-a)*b == -(a*b) exactly, so only the association of
// the tree decides the result, and the tree is kept as KPP builds it.
// ---------------------------------------------------------------------------
struct Node {
  enum K { CONST, ELM, ADD, SUB, MUL, DIV } k;
  int sign = 1;
  float cval = 0;     // CONST: absolute value
  std::string ref;    // ELM: printed reference, e.g. "V[3]"
  Node *l = nullptr, *r = nullptr;
};
std::vector<Node*> g_nodes;
Node* mk(Node::K k) {
  Node* n = new Node();
  n->k = k;
  g_nodes.push_back(n);
  return n;
}
Node* Const(double v) {
  Node* n = mk(Node::CONST);
  n->cval = (float)v;
  if (n->cval < 0) {
    n->cval = -n->cval;
    n->sign = -1;
  }
  return n;
}
Node* Ref(const std::string& s) {
  Node* n = mk(Node::ELM);
  n->ref = s;
  return n;
}
bool is_const(Node* n, float v) { return n && n->k == Node::CONST && n->cval == v; }
Node* bin(Node::K k, Node* a, Node* b) {
  Node* n = mk(k);
  n->l = a;
  n->r = b;
  return n;
}
Node* Add(Node* a, Node* b) {
  if (!a) return b;
  if (!b) return a;
  if (is_const(a, 0)) return b;
  if (is_const(b, 0)) return a;
  return bin(Node::ADD, a, b);
}
Node* Sub(Node* a, Node* b) {
  if (!a) return bin(Node::SUB, nullptr, b);
  if (!b) return a;
  if (is_const(a, 0)) return bin(Node::SUB, nullptr, b);
  if (is_const(b, 0)) return a;
  return bin(Node::SUB, a, b);
}
Node* Mul(Node* a, Node* b) {
  if (!a) return b;
  if (!b) return a;
  if (is_const(a, 1)) {  // the constant's sign moves to the other operand
    b->sign *= a->sign;
    return b;
  }
  if (is_const(b, 1)) return a;  // KPP drops the sign of a right unit constant
  if (is_const(a, 0)) return a;
  if (is_const(b, 0)) return b;
  return bin(Node::MUL, a, b);
}
Node* Div(Node* a, Node* b) {
  if (!a) return bin(Node::DIV, Const(1), b);
  if (!b) return a;
  if (is_const(b, 1)) return a;
  return bin(Node::DIV, a, b);
}

std::string cfmt(float v) {  // "%g" of the float, as src/code_c.c:66
  char b[64];
  snprintf(b, sizeof b, "%g", (double)v);
  std::string s = b;
  // make it a double literal in C/HIP without changing its value
  if (s.find_first_of(".eEn") == std::string::npos) s += ".0";
  return s;
}

// CFACTOR is written with "%12.6e" by KPP (src/gen.c:104-115 ascid).
std::string efmt(double v) {
  char b[64];
  snprintf(b, sizeof b, "%.6e", v);
  return b;
}

// Operation counter for the emitted arithmetic.
struct Ops {
  long add = 0, mul = 0, div = 0;
  long total() const { return add + mul + div; }
};

// Print a node. Leaf signs are folded into the enclosing +/- or into a
// leading unary minus; product/quotient chains print left-deep, which C
// evaluates left to right.
struct Printer {
  Ops* ops;
  int prec(Node* n) {
    switch (n->k) {
      case Node::ADD: case Node::SUB: return 1;
      case Node::MUL: case Node::DIV: return 2;
      default: return 3;
    }
  }
  std::string leaf(Node* n) { return n->k == Node::CONST ? cfmt(n->cval) : n->ref; }
  bool is_leaf(Node* n) { return n->k == Node::CONST || n->k == Node::ELM; }
  bool is_muldiv(Node* n) { return n->k == Node::MUL || n->k == Node::DIV; }
  // sign of a multiplicative chain: leaf signs hoisted out of the chain.
  // Signs stored on MUL/ADD nodes themselves are not printed by KPP
  // (src/code.c:692-764 writes leaf signs only), so they are ignored here.
  int chain_sign(Node* n) {
    if (is_leaf(n)) return n->sign;
    if (is_muldiv(n))
      return ((is_leaf(n->l) || is_muldiv(n->l)) ? chain_sign(n->l) : 1) *
             (is_leaf(n->r) ? n->r->sign : 1);
    return 1;
  }
  std::string unsigned_mul(Node* n) {  // product/quotient chain without signs
    if (is_leaf(n)) return leaf(n);
    std::string L = (is_leaf(n->l) || is_muldiv(n->l)) ? unsigned_mul(n->l) : "(" + expr(n->l) + ")";
    std::string R = is_leaf(n->r) ? leaf(n->r) : "(" + expr(n->r) + ")";
    if (n->k == Node::MUL) ops->mul++; else ops->div++;
    return L + (n->k == Node::MUL ? "*" : "/") + R;
  }
  // a term of a sum: returns (sign, text)
  std::pair<int, std::string> term(Node* n) {
    if (n->k == Node::ADD || n->k == Node::SUB) return {1, "(" + expr(n) + ")"};
    return {chain_sign(n), unsigned_mul(n)};
  }
  std::string expr(Node* n) {
    if (!n) return "0.0";
    if (n->k == Node::ADD || n->k == Node::SUB) {
      std::string L;
      if (n->l) {
        if (n->l->k == Node::ADD || n->l->k == Node::SUB) {
          L = expr(n->l);
        } else {
          auto t = term(n->l);
          L = (t.first < 0 ? "-" : "") + t.second;
        }
      }
      auto t = term(n->r);
      int s = t.first * (n->k == Node::SUB ? -1 : 1);
      if (n->l) ops->add++;
      if (!n->l) return (s < 0 ? "-" : "") + t.second;
      return L + (s < 0 ? "-" : "+") + t.second;
    }
    auto t = term(n);
    return (t.first < 0 ? "-" : "") + t.second;
  }
};

// ---------------------------------------------------------------------------
// Fortran 90 rate code.  Mechanisms written for KPP's F90 back end (the MCM
// export, examples/mcm) give their rate expressions, and the code KPP
// inlines into Update_RCONST, in Fortran.  This translator turns that subset
// into C with Fortran's typing rules: a real literal without a D exponent or
// a kind suffix is default REAL (single precision -- KPP's F90 build,
// util/Makefile_f90:38-39, compiles with gfortran and no -fdefault-real-8),
// an operation on two single-precision operands is done in single precision,
// and an operand of kind dp promotes the operation to double.  Variables of
// the rate code are REAL(dp); INTEGER, PARAMETER constants are integers.
// ---------------------------------------------------------------------------
namespace f90 {

enum Ty { TINT = 0, TR4 = 1, TR8 = 2 };

struct Tok {
  char k;  // 'n' number, 'i' identifier, 'o' operator, 'e' end
  std::string s;
};

std::vector<Tok> lex(const std::string& src) {
  std::vector<Tok> t;
  size_t i = 0, n = src.size();
  auto dig = [&](size_t k) { return k < n && isdigit((unsigned char)src[k]); };
  while (i < n) {
    char c = src[i];
    if (isspace((unsigned char)c)) { i++; continue; }
    if (dig(i) || (c == '.' && dig(i + 1))) {
      size_t b = i;
      while (dig(i)) i++;
      if (i < n && src[i] == '.') {
        i++;
        while (dig(i)) i++;
      }
      if (i < n && strchr("EeDd", src[i]) &&
          (dig(i + 1) || ((src[i + 1] == '+' || src[i + 1] == '-') && dig(i + 2)))) {
        i += 2;
        while (dig(i)) i++;
      }
      if (i < n && src[i] == '_') {
        i++;
        while (i < n && (isalnum((unsigned char)src[i]) || src[i] == '_')) i++;
      }
      t.push_back({'n', src.substr(b, i - b)});
      continue;
    }
    if (isalpha((unsigned char)c) || c == '_') {
      size_t b = i;
      while (i < n && (isalnum((unsigned char)src[i]) || src[i] == '_')) i++;
      t.push_back({'i', upper(src.substr(b, i - b))});
      continue;
    }
    if (c == '*' && i + 1 < n && src[i + 1] == '*') {
      t.push_back({'o', "**"});
      i += 2;
      continue;
    }
    if (strchr("+-*/(),=", c)) {
      t.push_back({'o', std::string(1, c)});
      i++;
      continue;
    }
    die("F90 rate code: unexpected character '" + std::string(1, c) + "' in: " + src);
  }
  t.push_back({'e', ""});
  return t;
}

struct E {
  std::string c;  // C text
  Ty t;
};

// literal -> C text and type
E number(const std::string& s) {
  std::string u = upper(s), body = u;
  bool kind8 = false;
  size_t us = u.find('_');
  if (us != std::string::npos) {
    std::string kind = u.substr(us + 1);
    body = u.substr(0, us);
    if (kind == "DP" || kind == "8") kind8 = true;
    else if (kind != "4" && kind != "SP") die("F90: unknown kind suffix in " + s);
  }
  bool dexp = body.find('D') != std::string::npos;
  bool real = dexp || body.find('.') != std::string::npos || body.find('E') != std::string::npos;
  if (!real && !kind8) return {body, TINT};
  for (auto& ch : body)
    if (ch == 'D') ch = 'E';
  if (!real) body += ".";
  if (dexp || kind8) return {body, TR8};
  return {body + "f", TR4};
}

struct Ctx {
  std::map<std::string, long> iparam;            // INTEGER, PARAMETER constants
  std::map<std::string, int> arrays;             // array name -> declared size
  std::map<std::string, std::string> inputs;     // TEMP, SUN, ... -> C name
  std::map<std::string, int> species;            // species name -> index (KPP order)
  std::function<std::string(int)> conc;          // C text of species k's concentration
};

struct Parser {
  std::vector<Tok> t;
  size_t p = 0;
  Ctx* cx;
  std::set<std::string>* uses;
  std::string src;

  bool op(const char* o) const { return t[p].k == 'o' && t[p].s == o; }
  void expect(const char* o) {
    if (!op(o)) die(std::string("F90: expected '") + o + "' in: " + src);
    p++;
  }
  static E bin(const E& a, const std::string& o, const E& b) {
    return {"(" + a.c + " " + o + " " + b.c + ")", std::max(a.t, b.t)};
  }
  static std::string as(const E& a, Ty t) {
    if (a.t == t) return a.c;
    return std::string(t == TR8 ? "(double)" : (t == TR4 ? "(float)" : "(int)")) + a.c;
  }
  E power(const E& a, const E& b) {
    if (b.t == TINT) {
      // integer exponent: repeated multiplication (kpp_powi_*: binary powering)
      const char* f = a.t == TR8 ? "kpp_powi_d" : (a.t == TR4 ? "kpp_powi_f" : "kpp_powi_i");
      return {std::string(f) + "(" + a.c + ", " + b.c + ")", a.t};
    }
    Ty t = std::max(a.t, b.t);
    if (t == TR4) return {"powf(" + as(a, TR4) + ", " + b.c + ")", TR4};
    return {"pow(" + as(a, TR8) + ", " + as(b, TR8) + ")", TR8};
  }
  E expr() {
    E a;
    if (op("-") || op("+")) {
      std::string o = t[p++].s;
      E b = term();
      a = (o == "-") ? E{"(-" + b.c + ")", b.t} : b;
    } else {
      a = term();
    }
    while (op("+") || op("-")) {
      std::string o = t[p++].s;
      a = bin(a, o, term());
    }
    return a;
  }
  E term() {
    E a = factor();
    while (op("*") || op("/")) {
      std::string o = t[p++].s;
      a = bin(a, o, factor());
    }
    return a;
  }
  E factor() {
    E a = primary();
    if (op("**")) {
      p++;
      E b = factor();
      return power(a, b);
    }
    return a;
  }
  E intrinsic(const std::string& f, std::vector<E>& args) {
    static const std::set<std::string> one = {"EXP", "LOG",  "LOG10", "SQRT", "COS", "SIN",
                                              "TAN", "ACOS", "ASIN",  "ATAN", "TANH", "ABS"};
    if (one.count(f)) {
      if (args.size() != 1) die("F90: " + f + " takes one argument in: " + src);
      E a = args[0];
      if (a.t == TINT && f != "ABS") die("F90: integer argument to " + f + " in: " + src);
      std::string name = f == "ABS" ? (a.t == TINT ? "abs" : "fabs") : lower(f);
      if (a.t == TR4) name += "f";
      return {name + "(" + a.c + ")", a.t};
    }
    if (f == "DBLE" && args.size() == 1) return {"((double)" + args[0].c + ")", TR8};
    if (f == "REAL" && args.size() == 1) return {"((float)" + args[0].c + ")", TR4};
    if ((f == "MAX" || f == "MIN") && args.size() >= 2) {
      E a = args[0];
      for (size_t k = 1; k < args.size(); k++) {
        Ty ty = std::max(a.t, args[k].t);
        std::string x = as(a, ty), y = as(args[k], ty);
        a = {"((" + x + ") " + (f == "MAX" ? ">" : "<") + " (" + y + ") ? (" + x + ") : (" + y +
                 "))",
             ty};
      }
      return a;
    }
    die("F90: unknown function " + f + " in: " + src);
  }
  E primary() {
    const Tok k = t[p];
    if (k.k == 'n') {
      p++;
      return number(k.s);
    }
    if (op("(")) {
      p++;
      E a = expr();
      expect(")");
      return {"(" + a.c + ")", a.t};
    }
    if (k.k != 'i') die("F90: unexpected token '" + k.s + "' in: " + src);
    p++;
    if (k.s == "C" && op("(")) {  // concentration of a species: C(ind_NAME)
      p++;
      if (t[p].k != 'i' || t[p].s.rfind("IND_", 0) != 0) die("F90: C(...) needs ind_NAME: " + src);
      std::string sp = t[p].s.substr(4);
      p++;
      expect(")");
      auto it = cx->species.find(sp);
      if (it == cx->species.end()) die("F90: unknown species " + sp + " in: " + src);
      uses->insert("C(" + sp + ")");
      return {cx->conc(it->second), TR8};
    }
    if (op("(")) {
      p++;
      std::vector<E> args;
      if (!op(")")) {
        args.push_back(expr());
        while (op(",")) {
          p++;
          args.push_back(expr());
        }
      }
      expect(")");
      auto ar = cx->arrays.find(k.s);
      if (ar != cx->arrays.end()) {
        if (args.size() != 1 || args[0].t != TINT) die("F90: bad index of " + k.s + ": " + src);
        long ix = strtol(args[0].c.c_str(), nullptr, 10);
        if (std::to_string(ix) != args[0].c || ix < 1 || ix > ar->second)
          die("F90: index of " + k.s + " must be a constant in 1.." +
              std::to_string(ar->second) + ": " + src);
        uses->insert(k.s + "(" + args[0].c + ")");
        return {"f_" + k.s + "[" + args[0].c + "]", TR8};
      }
      return intrinsic(k.s, args);
    }
    auto ip = cx->iparam.find(k.s);
    if (ip != cx->iparam.end()) return {std::to_string(ip->second), TINT};
    uses->insert(k.s);
    auto in = cx->inputs.find(k.s);
    if (in != cx->inputs.end()) return {in->second, TR8};
    return {"f_" + k.s, TR8};
  }
  static std::string lower(std::string s) {
    for (auto& ch : s) ch = (char)tolower((unsigned char)ch);
    return s;
  }
};

E translate(const std::string& src, Ctx& cx, std::set<std::string>* uses) {
  Parser ps{lex(src), 0, &cx, uses, src};
  E e = ps.expr();
  if (ps.t[ps.p].k != 'e') die("F90: trailing text in: " + src);
  return e;
}

struct Stmt {
  std::string target;  // "NAME" or "J(4)"
  std::string ctarget;
  E rhs;
  std::set<std::string> uses;
};

// Statement text -> lines: '!' comments removed, '&' continuations joined.
std::vector<std::string> logical_lines(const std::string& code) {
  std::vector<std::string> out;
  std::string cur;
  std::istringstream is(code);
  std::string line;
  bool cont = false;
  while (std::getline(is, line)) {
    size_t bang = line.find('!');
    if (bang != std::string::npos) line = line.substr(0, bang);
    size_t b = line.find_first_not_of(" \t\r");
    if (b == std::string::npos) continue;
    line = line.substr(b);
    if (cont && !line.empty() && line[0] == '&') line = line.substr(1);
    size_t e = line.find_last_not_of(" \t\r");
    line = line.substr(0, e + 1);
    bool more = !line.empty() && line.back() == '&';
    if (more) line.pop_back();
    cur += line + " ";
    cont = more;
    if (!more) {
      out.push_back(cur);
      cur.clear();
    }
  }
  if (!cur.empty()) out.push_back(cur);
  return out;
}

struct Program {
  Ctx cx;
  std::map<std::string, std::vector<std::string>> subs;  // SUBROUTINE -> statement lines
  std::vector<Stmt> stmts;

  // INTEGER, PARAMETER and array declarations, and subroutine bodies
  void module(const std::string& text) {
    std::string in_sub;
    auto word = [](const std::string& l, size_t from) {
      size_t b = l.find_first_not_of(" \t", from), e = b;
      while (e < l.size() && (isalnum((unsigned char)l[e]) || l[e] == '_')) e++;
      return std::make_pair(upper(b == std::string::npos ? "" : l.substr(b, e - b)), e);
    };
    for (auto& l : logical_lines(text)) {
      std::string u = upper(l);
      auto [w0, e0] = word(l, 0);
      if (w0.empty()) continue;
      if (!in_sub.empty()) {
        if (w0 == "END") {
          in_sub.clear();
          continue;
        }
        subs[in_sub].push_back(l);
        continue;
      }
      if (w0 == "SUBROUTINE") {
        in_sub = word(l, e0).first;
        subs[in_sub];
        continue;
      }
      size_t dc = u.find("::");
      if (dc == std::string::npos) continue;
      std::string decl = u.substr(0, dc), names = l.substr(dc + 2);
      if (decl.find("INTEGER") != std::string::npos && decl.find("PARAMETER") != std::string::npos) {
        auto tk = lex(names);
        for (size_t k = 0; k + 2 < tk.size(); k++)
          if (tk[k].k == 'i' && tk[k + 1].s == "=") {
            size_t q = k + 2;
            bool neg = false;
            if (tk[q].s == "-") { neg = true; q++; }
            if (tk[q].k != 'n') die("F90: INTEGER PARAMETER needs a literal: " + l);
            cx.iparam[tk[k].s] = (neg ? -1 : 1) * strtol(tk[q].s.c_str(), nullptr, 10);
          }
      } else if (decl.find("DIMENSION") != std::string::npos) {
        size_t a = decl.find("DIMENSION");
        auto tk = lex(decl.substr(a));
        if (tk.size() < 4 || tk[2].k != 'n') die("F90: DIMENSION needs a literal size: " + l);
        int size = atoi(tk[2].s.c_str());
        for (auto& x : lex(names))
          if (x.k == 'i') cx.arrays[x.s] = size;
      }
    }
  }

  void statement(const std::string& line) {
    auto tk = lex(line);
    if (tk[0].k != 'i') die("F90: cannot translate statement: " + line);
    const std::string& w0 = tk[0].s;
    if (w0 == "USE" || w0 == "IMPLICIT") return;
    if (w0 == "CALL") {
      if (tk[1].k != 'i') die("F90: bad CALL: " + line);
      auto s = subs.find(tk[1].s);
      if (s == subs.end()) die("F90: CALL of unknown subroutine " + tk[1].s);
      for (auto& l : s->second) statement(l);
      return;
    }
    size_t eq = line.find('=');
    if (eq == std::string::npos) die("F90: cannot translate statement: " + line);
    std::string lhs = line.substr(0, eq), rhs = line.substr(eq + 1);
    Stmt st;
    auto lt = lex(lhs);
    if (lt.size() == 2 && lt[0].k == 'i') {
      st.target = lt[0].s;
      st.ctarget = "f_" + lt[0].s;
    } else {
      std::set<std::string> u;
      E e = translate(lhs, cx, &u);
      if (u.size() != 1 || e.c.rfind("f_", 0) != 0 || e.c.find('[') == std::string::npos)
        die("F90: unsupported assignment target: " + line);
      st.target = *u.begin();
      st.ctarget = e.c;
    }
    st.rhs = translate(rhs, cx, &st.uses);
    stmts.push_back(st);
  }

  void code(const std::string& text) {
    for (auto& l : logical_lines(text)) statement(l);
  }

  // inputs a name depends on, through the statements before position `upto`
  std::set<std::string> deps(const std::set<std::string>& uses, size_t upto) const {
    std::set<std::string> out, todo(uses.begin(), uses.end());
    std::set<std::pair<std::string, size_t>> seen;
    std::vector<std::pair<std::string, size_t>> stack;
    for (auto& u : uses) stack.push_back({u, upto});
    while (!stack.empty()) {
      auto [name, pos] = stack.back();
      stack.pop_back();
      if (!seen.insert({name, pos}).second) continue;
      if (name.rfind("C(", 0) == 0) { out.insert("STATE"); continue; }
      if (cx.inputs.count(name)) { out.insert(name); continue; }
      size_t k = pos;
      bool found = false;
      while (k > 0) {
        k--;
        if (stmts[k].target == name) { found = true; break; }
      }
      if (!found) die("F90: '" + name + "' is used before it is assigned");
      for (auto& u : stmts[k].uses) stack.push_back({u, k});
    }
    return out;
  }

  // statements needed to evaluate expressions using `uses` (liveness, in order)
  std::vector<size_t> needed(std::set<std::string> live) const {
    std::vector<size_t> keep;
    for (size_t k = stmts.size(); k-- > 0;) {
      if (!live.count(stmts[k].target)) continue;
      keep.push_back(k);
      live.erase(stmts[k].target);
      for (auto& u : stmts[k].uses) live.insert(u);
    }
    std::reverse(keep.begin(), keep.end());
    return keep;
  }

  void emit(std::ostream& o, const std::vector<size_t>& keep, const char* ind = "  ") const {
    std::set<std::string> declared;
    for (auto& a : cx.arrays) {
      bool used = false;
      for (size_t k : keep)
        if (stmts[k].target.rfind(a.first + "(", 0) == 0) used = true;
      if (used) o << ind << "double f_" << a.first << "[" << a.second + 1 << "];\n";
    }
    for (size_t k : keep) {
      const Stmt& s = stmts[k];
      bool scalar = s.ctarget.find('[') == std::string::npos;
      if (scalar && declared.insert(s.ctarget).second)
        o << ind << "double " << s.ctarget << " = " << s.rhs.c << ";\n";
      else
        o << ind << s.ctarget << " = " << s.rhs.c << ";\n";
    }
  }
};

const char* kPowiHelpers = R"(
KPPQ double kpp_powi_d(double x, int n) {
  double r = 1.0; unsigned m = n < 0 ? -(unsigned)n : (unsigned)n;
  for (; m; m >>= 1) { if (m & 1u) r *= x; x *= x; }
  return n < 0 ? 1.0 / r : r;
}
KPPQ float kpp_powi_f(float x, int n) {
  float r = 1.0f; unsigned m = n < 0 ? -(unsigned)n : (unsigned)n;
  for (; m; m >>= 1) { if (m & 1u) r *= x; x *= x; }
  return n < 0 ? 1.0f / r : r;
}
KPPQ int kpp_powi_i(int x, int n) {
  int r = 1;
  if (n < 0) return (x == 1) ? 1 : ((x == -1) ? ((n & 1) ? -1 : 1) : 0);
  for (; n; n >>= 1) { if (n & 1) r *= x; x *= x; }
  return r;
}
)";

}  // namespace f90

// ---------------------------------------------------------------------------
// Model code construction shared by both emitters.
// ---------------------------------------------------------------------------
struct Model {
  Mech* m;
  Structure S;
  // A[j] for used reactions
  std::vector<int> a_used;
  std::vector<std::string> a_expr;
  std::vector<std::string> vdot_expr;
  // B terms: (eqn i, var j) -> name, expression
  std::map<std::pair<int, int>, int> b_index;
  std::vector<std::string> b_expr;
  std::vector<std::string> jvs_expr;  // per LU entry
  std::vector<char> jvs_zero;
  // Fun_SPLIT (#FUNCTION SPLIT): A[j] of the reactions producing a variable,
  // production P[i] and destruction frequency D[i]
  bool fun_split = false;
  // the adjoint's model functions (src/gen.c:1327-1600, 1770-1870, 2035-2106)
  bool adjoint = false;
  std::vector<std::string> d2a_expr, hess_expr, htu_expr, juv_expr, jtuv_expr;
  std::vector<int> ih_i, ih_j, ih_k;
  Ops hess_ops;
  void build_adjoint();
  std::vector<int> as_used;
  std::vector<std::string> as_expr, p_expr, d_expr;
  Ops fun_ops, jac_ops, lu_ops, sol_ops, split_ops;
  std::vector<std::pair<int, std::string>> rconst;  // (reaction, C expression)
  // rate classes: 0 = fixed per cell, 1 = SUN/TIME, 2 = state (concentrations)
  std::map<int, int> rate_class;
  f90::Program* f90 = nullptr;                     // F90 rate code, or null (C syntax)
  std::map<int, f90::E> f90_rate;                  // translated rate expressions
  std::map<int, std::set<std::string>> f90_uses;

  // C text of rate j's expression
  std::string rate_c(int j) const {
    if (f90) return f90_rate.at(j).c;
    return m->rates[j].expr;
  }
  void classify();
  // statements of the F90 preamble needed by the given rates
  std::vector<size_t> preamble(const std::vector<int>& rates) const {
    std::set<std::string> live;
    if (!f90) return {};
    for (int j : rates)
      for (auto& u : f90_uses.at(j)) live.insert(u);
    return f90->needed(live);
  }

  Node* rconst_node(int j) {
    auto& r = m->rates[j];
    if (r.type == NUMBER) return Const(r.f);
    return Ref("RCT[" + std::to_string(j) + "]");
  }

  void build() {
    int nv = S.nvar, ns = S.nspc, ne = S.nreact;
    Printer pf{&fun_ops};
    // Fun (src/gen.c:713-761): A[j] for reactions touching a variable.
    for (int j = 0; j < ne; j++) {
      bool used = false;
      for (int i = 0; i < nv; i++)
        if (m->st[i][j] != 0) { used = true; break; }
      if (!used) continue;
      Node* p = rconst_node(j);
      int i = 0;
      for (; i < nv; i++)
        for (int k = 1; k <= (int)m->sl[i][j]; k++) p = Mul(p, Ref("V[" + std::to_string(i) + "]"));
      for (; i < ns; i++)
        for (int k = 1; k <= (int)m->sl[i][j]; k++)
          p = Mul(p, Ref("F[" + std::to_string(i - nv) + "]"));
      a_used.push_back(j);
      a_expr.push_back(pf.expr(p));
    }
    for (int i = 0; i < nv; i++) {
      Node* sum = Const(0);
      for (int j = 0; j < ne; j++)
        sum = Add(sum, Mul(Const(m->st[i][j]), Ref("A[" + std::to_string(j) + "]")));
      vdot_expr.push_back(pf.expr(sum));
    }
    if (fun_split) {
      // Fun_SPLIT (src/gen.c:713-812, z_useAggregate = 0): A[j] for the
      // reactions with a variable product; P = sum Stoich_Right*A; D = sum
      // over the reactions consuming i of RCT*Stoich_Left times the other
      // reactants (the loss frequency); Vdot = P - D*V
      Printer ps{&split_ops};
      for (int j = 0; j < ne; j++) {
        bool used = false;
        for (int i = 0; i < nv; i++)
          if (m->sr[i][j] != 0) { used = true; break; }
        if (!used) continue;
        Node* p = rconst_node(j);
        int i = 0;
        for (; i < nv; i++)
          for (int k = 1; k <= (int)m->sl[i][j]; k++) p = Mul(p, Ref("V[" + std::to_string(i) + "]"));
        for (; i < ns; i++)
          for (int k = 1; k <= (int)m->sl[i][j]; k++)
            p = Mul(p, Ref("F[" + std::to_string(i - nv) + "]"));
        as_used.push_back(j);
        as_expr.push_back(ps.expr(p));
      }
      for (int i = 0; i < nv; i++) {
        Node* sum = Const(0);
        for (int j = 0; j < ne; j++)
          sum = Add(sum, Mul(Const(m->sr[i][j]), Ref("A[" + std::to_string(j) + "]")));
        p_expr.push_back(ps.expr(sum));
      }
      for (int i = 0; i < nv; i++) {
        Node* sum = Const(0);
        for (int j = 0; j < ne; j++) {
          if (m->sl[i][j] == 0) continue;
          Node* p = Mul(rconst_node(j), Const(m->sl[i][j]));
          int l = 0;
          for (; l < nv; l++) {
            int mm = (int)m->sl[l][j] - (l == i);
            for (int k = 1; k <= mm; k++) p = Mul(p, Ref("V[" + std::to_string(l) + "]"));
          }
          for (; l < ns; l++)
            for (int k = 1; k <= (int)m->sl[l][j]; k++)
              p = Mul(p, Ref("F[" + std::to_string(l - nv) + "]"));
          sum = Add(sum, p);
        }
        d_expr.push_back(ps.expr(sum));
      }
    }
    // Jac_SP (src/gen.c:1239-1284)
    Printer pj{&jac_ops};
    for (int i = 0; i < ne; i++)
      for (int j = 0; j < nv; j++) {
        if (m->sl[j][i] == 0) continue;
        Node* p = Mul(rconst_node(i), Const(m->sl[j][i]));
        int l = 0;
        for (; l < nv; l++) {
          int mm = (int)m->sl[l][i] - (l == j);
          for (int k = 1; k <= mm; k++) p = Mul(p, Ref("V[" + std::to_string(l) + "]"));
        }
        for (; l < ns; l++)
          for (int k = 1; k <= (int)m->sl[l][i]; k++)
            p = Mul(p, Ref("F[" + std::to_string(l - nv) + "]"));
        b_index[{i, j}] = (int)b_expr.size();
        b_expr.push_back(pj.expr(p));
      }
    for (int i = 0; i < nv; i++)
      for (int j = 0; j < nv; j++) {
        if (!S.LU[i][j]) continue;
        Node* sum = Const(0);
        for (int k = 0; k < ne; k++) {
          if (m->st[i][k] != 0 && m->sl[j][k] != 0)
            sum = Add(sum, Mul(Const(m->st[i][k]),
                               Ref("B[" + std::to_string(b_index[{k, j}]) + "]")));
        }
        bool z = is_const(sum, 0);
        jvs_zero.push_back(z);
        jvs_expr.push_back(z ? std::string("0.0") : pj.expr(sum));
      }
    // op counts of the LU (src/util sutil.c:2-23 arithmetic) and the solve
    for (int k = 0; k < nv; k++)
      for (int kk = S.lu_crow[k]; kk < S.lu_diag[k]; kk++) {
        int j = S.lu_icol[kk];
        lu_ops.div++;
        for (int jj = S.lu_diag[j] + 1; jj < S.lu_crow[j + 1]; jj++) {
          lu_ops.mul++;
          lu_ops.add++;
        }
      }
    for (int i = 0; i < nv; i++) {
      sol_ops.mul += (S.lu_diag[i] - S.lu_crow[i]) + (S.lu_crow[i + 1] - S.lu_diag[i] - 1);
      sol_ops.add += (S.lu_diag[i] - S.lu_crow[i]) + (S.lu_crow[i + 1] - S.lu_diag[i] - 1);
      sol_ops.div += 1;
    }
    for (int j = 0; j < ne; j++)
      if (m->rates[j].type != NUMBER) rconst.push_back({j, m->rates[j].expr});
  }
};

// The model functions of the discrete adjoint (int/rosenbrock_adj.c).
//  Hessian (src/gen.c:1327-1498): D2A[k] = d2 A_j/dV_i1 dV_i2 (i1 <= i2) for
//   every reaction with the reactants, then HESS[m] = d2 f_i/dV_j dV_k =
//   sum_j Stoich[i][j] D2A, entries (i, j <= k) in that loop order;
//  HessTR_Vec (:1527-1549): HTU_i = sum over entries with j == i of
//   HESS*(U1[i_]*U2[k]), and with k == i of HESS*(U1[i_]*U2[j]);
//  Jac_SP_Vec / JacTR_SP_Vec (src/gen.c:1770-1870): J*U and J^T*U over the
//   Jacobian's structural entries (structJ) stored in the LU pattern.
void Model::build_adjoint() {
  const int nv = S.nvar, ns = S.nspc, ne = S.nreact;
  Printer ph{&hess_ops};
  std::map<std::tuple<int, int, int>, int> d2a_of;  // (j, i1, i2) -> D2A index
  auto V = [](int i) { return Ref("V[" + std::to_string(i) + "]"); };
  auto Fx = [&](int i) { return Ref("F[" + std::to_string(i - nv) + "]"); };
  for (int j = 0; j < ne; j++)
    for (int i1 = 0; i1 < nv; i1++)
      for (int i2 = i1; i2 < nv; i2++) {
        Node* p = nullptr;
        if (i1 == i2) {
          if (m->sl[i1][j] < 2) continue;
          p = rconst_node(j);
          for (int i = 0; i < i1; i++)
            for (int k = 1; k <= (int)m->sl[i][j]; k++) p = Mul(p, V(i));
          p = Mul(p, Const(m->sl[i1][j]));
          p = Mul(p, Const(m->sl[i1][j] - 1));
          for (int k = 1; k <= (int)m->sl[i1][j] - 2; k++) p = Mul(p, V(i1));
          int i = i1 + 1;
          for (; i < nv; i++)
            for (int k = 1; k <= (int)m->sl[i][j]; k++) p = Mul(p, V(i));
          for (; i < ns; i++)
            for (int k = 1; k <= (int)m->sl[i][j]; k++) p = Mul(p, Fx(i));
        } else {
          if (!(m->sl[i1][j] >= 1 && m->sl[i2][j] >= 1)) continue;
          p = rconst_node(j);
          for (int i = 0; i < i1; i++)
            for (int k = 1; k <= (int)m->sl[i][j]; k++) p = Mul(p, V(i));
          p = Mul(p, Const(m->sl[i1][j]));
          for (int k = 1; k <= (int)m->sl[i1][j] - 1; k++) p = Mul(p, V(i1));
          for (int i = i1 + 1; i < i2; i++)
            for (int k = 1; k <= (int)m->sl[i][j]; k++) p = Mul(p, V(i));
          p = Mul(p, Const(m->sl[i2][j]));
          for (int k = 1; k <= (int)m->sl[i2][j] - 1; k++) p = Mul(p, V(i2));
          int i = i2 + 1;
          for (; i < nv; i++)
            for (int k = 1; k <= (int)m->sl[i][j]; k++) p = Mul(p, V(i));
          for (; i < ns; i++)
            for (int k = 1; k <= (int)m->sl[i][j]; k++) p = Mul(p, Fx(i));
        }
        d2a_of[{j, i1, i2}] = (int)d2a_expr.size();
        d2a_expr.push_back(ph.expr(p));
      }
  // reactions per (i1, i2) pair, in j order
  std::map<std::pair<int, int>, std::vector<std::pair<int, int>>> by_pair;
  for (auto& kv : d2a_of)
    by_pair[{std::get<1>(kv.first), std::get<2>(kv.first)}].push_back(
        {std::get<0>(kv.first), kv.second});
  for (auto& kv : by_pair) std::sort(kv.second.begin(), kv.second.end());
  for (int i = 0; i < nv; i++)
    for (auto& kv : by_pair) {  // (i1, i2) in lexicographic order, as the loops
      Node* sum = Const(0);
      bool any = false;
      for (auto& jk : kv.second)
        if (m->st[i][jk.first] != 0) {
          sum = Add(sum, Mul(Const(m->st[i][jk.first]), Ref("D2A[" + std::to_string(jk.second) + "]")));
          any = true;
        }
      if (!any) continue;
      ih_i.push_back(i);
      ih_j.push_back(kv.first.first);
      ih_k.push_back(kv.first.second);
      hess_expr.push_back(ph.expr(sum));
    }
  const int nh = (int)hess_expr.size();
  for (int i = 0; i < nv; i++) {
    Node* sum = Const(0);
    for (int k = 0; k < nh; k++) {
      const std::string h = "HESS[" + std::to_string(k) + "]";
      if (ih_j[k] == i)
        sum = Add(sum, Mul(Ref(h), Mul(Ref("U1[" + std::to_string(ih_i[k]) + "]"),
                                       Ref("U2[" + std::to_string(ih_k[k]) + "]"))));
      else if (ih_k[k] == i)
        sum = Add(sum, Mul(Ref(h), Mul(Ref("U1[" + std::to_string(ih_i[k]) + "]"),
                                       Ref("U2[" + std::to_string(ih_j[k]) + "]"))));
    }
    htu_expr.push_back(ph.expr(sum));
  }
  std::map<std::pair<int, int>, int> lu_at;
  for (int r = 0; r < nv; r++)
    for (int q = S.lu_crow[r]; q < S.lu_crow[r + 1]; q++) lu_at[{r, S.lu_icol[q]}] = q;
  for (int i = 0; i < nv; i++) {
    Node* sum = Const(0);
    for (int j = 0; j < nv; j++)
      if (S.LU[i][j] && S.J[i][j] != 0)
        sum = Add(sum, Mul(Ref("JVS[" + std::to_string(lu_at[{i, j}]) + "]"),
                           Ref("UV[" + std::to_string(j) + "]")));
    juv_expr.push_back(ph.expr(sum));
  }
  for (int i = 0; i < nv; i++) {
    Node* sum = Const(0);
    for (int j = 0; j < nv; j++)
      if (S.J[j][i] != 0)
        sum = Add(sum, Mul(Ref("JVS[" + std::to_string(lu_at[{j, i}]) + "]"),
                           Ref("UV[" + std::to_string(j) + "]")));
    jtuv_expr.push_back(ph.expr(sum));
  }
}

// KppSolveTR (src/gen.c:2035-2106): (LU)^T x = b, U^T forward then L^T
// backward; X may alias XX.  Statements per row.
std::vector<std::string> solve_tr_stmts(const Structure& S) {
  const int nv = S.nvar;
  std::map<std::pair<int, int>, int> at;  // (row, col) -> LU index
  for (int r = 0; r < nv; r++)
    for (int q = S.lu_crow[r]; q < S.lu_crow[r + 1]; q++) at[{r, S.lu_icol[q]}] = q;
  std::vector<std::string> st;
  Ops dummy;
  Printer pr{&dummy};
  for (int i = 0; i < nv; i++) {  // XX_i = (X_i - sum_{j<i} U_ji XX_j) / U_ii
    Node* sum = Ref("X[" + std::to_string(i) + "]");
    for (int j = 0; j < i; j++) {
      auto it = at.find({j, i});
      if (it != at.end())
        sum = Sub(sum, Mul(Ref("JVS[" + std::to_string(it->second) + "]"),
                           Ref("XX[" + std::to_string(j) + "]")));
    }
    sum = Div(sum, Ref("JVS[" + std::to_string(S.lu_diag[i]) + "]"));
    st.push_back("  XX[" + std::to_string(i) + "] = " + pr.expr(sum) + ";\n");
  }
  for (int i = nv - 1; i >= 0; i--) {  // XX_i -= sum_{j>i} L_ji XX_j
    Node* sum = Ref("XX[" + std::to_string(i) + "]");
    bool any = false;
    for (int j = i + 1; j < nv; j++) {
      auto it = at.find({j, i});
      if (it != at.end()) {
        sum = Sub(sum, Mul(Ref("JVS[" + std::to_string(it->second) + "]"),
                           Ref("XX[" + std::to_string(j) + "]")));
        any = true;
      }
    }
    if (any) st.push_back("  XX[" + std::to_string(i) + "] = " + pr.expr(sum) + ";\n");
  }
  return st;
}

bool mentions(const std::string& expr, const std::string& ident) {
  size_t p = 0;
  while ((p = expr.find(ident, p)) != std::string::npos) {
    bool lb = (p == 0) || !(isalnum((unsigned char)expr[p - 1]) || expr[p - 1] == '_');
    size_t e = p + ident.size();
    bool rb = (e >= expr.size()) || !(isalnum((unsigned char)expr[e]) || expr[e] == '_');
    if (lb && rb) return true;
    p = e;
  }
  return false;
}

void Model::classify() {
  for (auto& pr : rconst) {
    const int j = pr.first;
    if (f90) {
      f90_rate[j] = f90::translate(m->rates[j].expr, f90->cx, &f90_uses[j]);
      auto d = f90->deps(f90_uses[j], f90->stmts.size());
      rate_class[j] = d.count("STATE") ? 2 : ((d.count("SUN") || d.count("TIME")) ? 1 : 0);
    } else {
      const std::string& e = pr.second;
      if (e.find("**") != std::string::npos)
        die("rate of reaction " + std::to_string(j + 1) + " uses '**': pass --rates-f90");
      rate_class[j] = (mentions(e, "SUN") || mentions(e, "TIME")) ? 1 : 0;
    }
  }
}

// Rate laws of util/UserRateLaws.c:9-78 (parameters are float, arithmetic
// double), written out for the emitted code.  The identifiers TEMP and
// CFACTOR are the enclosing function's per-cell values.
const char* kRateLawsBody = R"(
KPPQ double kpp_ARR_abc(float A0, float B0, float C0, double TEMP) {
  return (double)A0 * exp(-(double)B0 / TEMP) * pow((TEMP / 300.0), (double)C0);
}
KPPQ double kpp_ARR_ab(float A0, float B0, double TEMP) {
  return (double)A0 * exp(-(double)B0 / TEMP);
}
KPPQ double kpp_ARR_ac(float A0, float C0, double TEMP) {
  return (double)A0 * pow((TEMP / 300.0), (double)C0);
}
KPPQ double kpp_EP2(float A0, float C0, float A2, float C2, float A3, float C3, double TEMP,
                    double CFACTOR) {
  double K0, K2, K3;
  K0 = (double)A0 * exp(-(double)C0 / TEMP);
  K2 = (double)A2 * exp(-(double)C2 / TEMP);
  K3 = (double)A3 * exp(-(double)C3 / TEMP);
  K3 = K3 * CFACTOR * 1.0e+6;
  return K0 + K3 / (1.0 + K3 / K2);
}
KPPQ double kpp_EP3(float A1, float C1, float A2, float C2, double TEMP, double CFACTOR) {
  double K1, K2;
  K1 = (double)A1 * exp(-(double)C1 / TEMP);
  K2 = (double)A2 * exp(-(double)C2 / TEMP);
  return K1 + K2 * (1.0e+6 * CFACTOR);
}
KPPQ double kpp_FALL(float A0, float B0, float C0, float A1, float B1, float C1, float CF,
                     double TEMP, double CFACTOR) {
  double K0, K1;
  K0 = (double)A0 * exp(-(double)B0 / TEMP) * pow((TEMP / 300.0), (double)C0);
  K1 = (double)A1 * exp(-(double)B1 / TEMP) * pow((TEMP / 300.0), (double)C1);
  K0 = K0 * CFACTOR * 1.0e+6;
  K1 = K0 / K1;
  return (K0 / (1.0 + K1)) * pow((double)CF, (1.0 / (1.0 + pow((log10(K1)), 2))));
}
#define ARR_abc(a, b, c) kpp_ARR_abc((a), (b), (c), TEMP)
#define ARR_ab(a, b) kpp_ARR_ab((a), (b), TEMP)
#define ARR_ac(a, c) kpp_ARR_ac((a), (c), TEMP)
#define EP2(a0, c0, a2, c2, a3, c3) kpp_EP2((a0), (c0), (a2), (c2), (a3), (c3), TEMP, CFACTOR)
#define EP3(a1, c1, a2, c2) kpp_EP3((a1), (c1), (a2), (c2), TEMP, CFACTOR)
#define FALL(a0, b0, c0, a1, b1, c1, cf) \
  kpp_FALL((a0), (b0), (c0), (a1), (b1), (c1), (cf), TEMP, CFACTOR)
)";

std::string ident_ok(const std::string& root) {
  std::string s;
  for (char c : root) s += (isalnum((unsigned char)c) ? c : '_');
  return s;
}

void emit_int_array(std::ostream& o, const char* qual, const std::string& name,
                    const std::vector<int>& v) {
  o << qual << " int " << name << "[" << std::max<size_t>(v.size(), 1) << "] = {";
  for (size_t i = 0; i < v.size(); i++) {
    if (i % 16 == 0) o << "\n  ";
    o << v[i] << (i + 1 < v.size() ? "," : "");
  }
  if (v.empty()) o << "0";
  o << "};\n";
}

// initial value of each species in the final order, or the group default
std::vector<std::string> init_texts(Model& md) {
  std::vector<std::string> out;
  for (int k = 0; k < md.S.nspc; k++) {
    auto& sp = md.m->spc(md.m->code[k]);
    std::string v = sp.ival;
    if (v.empty()) v = (k < md.S.nvar) ? md.m->varDefault : md.m->fixDefault;
    out.push_back(v);
  }
  return out;
}

void emit_c_model(Model& md, const std::string& path) {
  Mech& m = *md.m;
  auto& S = md.S;
  std::ofstream o(path);
  o << "/* Generated by kppgen from the KPP mechanism '" << m.root << "'.\n"
    << "   Plain C model functions with the semantics of KPP's C output\n"
    << "   (src/gen.c GenerateFun/GenerateJac/GenerateSolve/GenerateUpdateRconst). */\n";
  o << "#ifndef KPPGEN_" << ident_ok(m.root) << "_MODEL_H\n#define KPPGEN_" << ident_ok(m.root)
    << "_MODEL_H\n#include <math.h>\n\n";
  o << "#define KPP_ROOT \"" << m.root << "\"\n";
  o << "#define KPP_NVAR " << S.nvar << "\n#define KPP_NFIX " << S.nfix << "\n#define KPP_NSPEC "
    << S.nspc << "\n#define KPP_NREACT " << S.nreact << "\n#define KPP_NONZERO " << S.nonzero
    << "\n#define KPP_LU_NONZERO " << S.lu_nonzero << "\n#define KPP_CFACTOR " << efmt(m.cfactor)
    << "\n\n";
  emit_int_array(o, "static const", "kpp_LU_ICOL", S.lu_icol);
  emit_int_array(o, "static const", "kpp_LU_IROW", S.lu_irow);
  emit_int_array(o, "static const", "kpp_LU_CROW", S.lu_crow);
  emit_int_array(o, "static const", "kpp_LU_DIAG", S.lu_diag);
  o << "static const char* const kpp_SPC_NAMES[" << S.nspc << "] = {";
  for (int k = 0; k < S.nspc; k++) o << (k % 8 ? " " : "\n  ") << "\"" << m.spc(m.code[k]).name << "\",";
  o << "};\n";
  auto iv = init_texts(md);
  o << "/* #INITVALUES text times CFACTOR, as Initialize() (src/gen.c:3252-3290) */\n";
  o << "static void kpp_Initialize(double* C) {\n  const double CFACTOR = " << efmt(m.cfactor)
    << ";\n";
  for (int k = 0; k < S.nspc; k++) o << "  C[" << k << "] = (" << iv[k] << ")*CFACTOR;\n";
  o << "}\n";
  o << "#define KPPQ static\n" << kRateLawsBody << "#undef KPPQ\n";
  o << "/* Update_RCONST (src/gen.c:2343-2473); NUMBER rates are folded into\n"
    << "   Fun/Jac_SP as constants, so their RCONST slot is only informative. */\n";
  o << "#define KPPQ static\n" << f90::kPowiHelpers << "#undef KPPQ\n";
  o << "/* V, F: the state the rates see (KPP's global C: the last accepted\n"
    << "   concentrations); only rate code that reads concentrations uses them. */\n";
  o << "static void kpp_Update_RCONST(double TEMP, double SUN, double CFACTOR, double TIME,\n"
    << "                              const double* V, const double* F, double* RCONST) {\n"
    << "  (void)SUN; (void)TIME; (void)TEMP; (void)CFACTOR; (void)V; (void)F;\n";
  {
    std::vector<int> all;
    for (auto& pr : md.rconst) all.push_back(pr.first);
    if (md.f90) md.f90->emit(o, md.preamble(all));
  }
  for (int j = 0; j < S.nreact; j++) {
    auto& r = m.rates[j];
    if (r.type == NUMBER) o << "  RCONST[" << j << "] = " << cfmt((float)r.f) << ";\n";
    else o << "  RCONST[" << j << "] = (" << md.rate_c(j) << ");\n";
  }
  o << "}\n";
  o << "static void kpp_Fun(const double* V, const double* F, const double* RCT, double* "
       "Vdot) {\n  double A[" << std::max(S.nreact, 1) << "];\n  (void)F; (void)RCT;\n";
  for (size_t q = 0; q < md.a_used.size(); q++)
    o << "  A[" << md.a_used[q] << "] = " << md.a_expr[q] << ";\n";
  for (int i = 0; i < S.nvar; i++) o << "  Vdot[" << i << "] = " << md.vdot_expr[i] << ";\n";
  o << "}\n";
  o << "#define KPP_FUN_SPLIT " << (md.fun_split ? 1 : 0) << "\n";
  if (md.fun_split) {
    o << "/* Fun_SPLIT (src/gen.c:713-812): Vdot = P - D*V */\n";
    o << "static void kpp_Fun_SPLIT(const double* V, const double* F, const double* RCT, "
         "double* Vdot) {\n  double A[" << std::max(S.nreact, 1) << "], P[" << S.nvar << "], D["
      << S.nvar << "];\n  (void)F; (void)RCT;\n";
    for (size_t q = 0; q < md.as_used.size(); q++)
      o << "  A[" << md.as_used[q] << "] = " << md.as_expr[q] << ";\n";
    for (int i = 0; i < S.nvar; i++) o << "  P[" << i << "] = " << md.p_expr[i] << ";\n";
    for (int i = 0; i < S.nvar; i++) o << "  D[" << i << "] = " << md.d_expr[i] << ";\n";
    o << "  for (int n = 0; n < " << S.nvar << "; n++) Vdot[n] = P[n] - D[n]*V[n];\n}\n";
  }
  o << "static void kpp_Jac_SP(const double* V, const double* F, const double* RCT, double* "
       "JVS) {\n  double B[" << std::max<size_t>(md.b_expr.size(), 1)
    << "];\n  (void)V; (void)F; (void)RCT;\n";
  for (size_t q = 0; q < md.b_expr.size(); q++) o << "  B[" << q << "] = " << md.b_expr[q] << ";\n";
  for (int n = 0; n < S.lu_nonzero; n++) o << "  JVS[" << n << "] = " << md.jvs_expr[n] << ";\n";
  o << "}\n";
  // KppSolve (src/gen.c:1960-2033)
  o << "static void kpp_KppSolve(const double* JVS, double* X) {\n";
  for (int i = 0; i < S.nvar; i++) {
    int b = S.lu_crow[i], e = S.lu_diag[i];
    if (b < e) {
      o << "  X[" << i << "] = X[" << i << "]";
      for (int j = b; j < e; j++) o << "-JVS[" << j << "]*X[" << S.lu_icol[j] << "]";
      o << ";\n";
    }
  }
  for (int i = S.nvar - 1; i >= 0; i--) {
    int b = S.lu_diag[i] + 1, e = S.lu_crow[i + 1];
    if (b < e) {
      o << "  X[" << i << "] = (X[" << i << "]";
      for (int j = b; j < e; j++) o << "-JVS[" << j << "]*X[" << S.lu_icol[j] << "]";
      o << ")/JVS[" << S.lu_diag[i] << "];\n";
    } else {
      o << "  X[" << i << "] = X[" << i << "]/JVS[" << S.lu_diag[i] << "];\n";
    }
  }
  o << "}\n";
  o << "#define KPP_ADJOINT " << (md.adjoint ? 1 : 0) << "\n";
  if (md.adjoint) {
    const size_t nh = md.hess_expr.size();
    o << "/* the discrete adjoint's model functions (src/gen.c:1327-1600, 1770-1870,\n"
         "   2035-2106): HESS(m) = d2 f_i/dV_j dV_k, i = IHESS_I(m), j <= k */\n";
    o << "#define KPP_NHESS " << std::max<size_t>(nh, 1) << "\n";
    auto arr = [&](const char* name, const std::vector<int>& v) {
      o << "static const int " << name << "[" << std::max<size_t>(v.size(), 1) << "] = {";
      for (size_t q = 0; q < v.size(); q++) o << (q ? "," : "") << v[q];
      if (v.empty()) o << "0";
      o << "};\n";
    };
    arr("kpp_IHESS_I", md.ih_i);
    arr("kpp_IHESS_J", md.ih_j);
    arr("kpp_IHESS_K", md.ih_k);
    o << "static void kpp_Hessian(const double* V, const double* F, const double* RCT, double* "
         "HESS) {\n  double D2A[" << std::max<size_t>(md.d2a_expr.size(), 1)
      << "];\n  (void)V; (void)F; (void)RCT; (void)D2A;\n";
    for (size_t q = 0; q < md.d2a_expr.size(); q++)
      o << "  D2A[" << q << "] = " << md.d2a_expr[q] << ";\n";
    for (size_t q = 0; q < nh; q++) o << "  HESS[" << q << "] = " << md.hess_expr[q] << ";\n";
    if (nh == 0) o << "  HESS[0] = 0.0;\n";
    o << "}\n";
    o << "static void kpp_HessTR_Vec(const double* HESS, const double* U1, const double* U2, "
         "double* HTU) {\n  (void)HESS; (void)U1; (void)U2;\n";
    for (int i = 0; i < S.nvar; i++) o << "  HTU[" << i << "] = " << md.htu_expr[i] << ";\n";
    o << "}\n";
    o << "static void kpp_Jac_SP_Vec(const double* JVS, const double* UV, double* JUV) {\n";
    for (int i = 0; i < S.nvar; i++) o << "  JUV[" << i << "] = " << md.juv_expr[i] << ";\n";
    o << "}\n";
    o << "static void kpp_JacTR_SP_Vec(const double* JVS, const double* UV, double* JTUV) {\n";
    for (int i = 0; i < S.nvar; i++) o << "  JTUV[" << i << "] = " << md.jtuv_expr[i] << ";\n";
    o << "}\n";
    o << "static void kpp_KppSolveTR(const double* JVS, const double* X, double* XX) {\n";
    for (auto& st : solve_tr_stmts(S)) o << st;
    o << "}\n";
  }
  o << "#undef ARR_abc\n#undef ARR_ab\n#undef ARR_ac\n#undef EP2\n#undef EP3\n#undef FALL\n";
  o << "#endif\n";
}

// ---------------------------------------------------------------------------
// Memory-pipelined emission of KppDecomp and KppSolve for the device.
//
// The routines are rewritten as a list of single-operation statements in the
// reference's order (one subtraction or division per statement, so the
// floating-point results are unchanged), over array elements with versions.
// Each operand is then taken either from a register (its producer is at most
// `window` statements back) or from memory through a load that is issued
// `dist` reads ahead of its use, but never before the store of the version it
// reads.  sched_barrier every `chunk` statements keeps the compiler from
// sinking the loads back next to their uses, so a wavefront keeps `dist`
// loads in flight instead of one.
// ---------------------------------------------------------------------------
struct PipeParams {
  int dist = 24, window = 48, chunk = 12;
};
PipeParams g_pipe;
bool g_emit_pipelined = false;  // --pipelined; the integrator does not use them
int g_lds_split = 0;   // LU entries the KppDecompL/KppSolveL variants may mirror in LDS (0: no variants)
long g_part_ops = 0;  // split routines into out-of-line parts of this many operations (0: no)

struct PipeOpd {
  int arr, idx, ver;
};
struct PipeStmt {
  int arr, idx;                // element written (new version)
  std::string expr;            // with $0, $1, ... for operands
  std::vector<PipeOpd> opds;
};

struct Pipeline {
  std::vector<std::string> arr_name;  // accessor names, e.g. "G", "X"
  std::vector<PipeStmt> st;
  std::map<std::pair<int, int>, int> cur;               // (arr, idx) -> current version
  std::map<std::tuple<int, int, int>, int> producer;    // version -> statement

  PipeOpd read(int a, int i) { return {a, i, cur[{a, i}]}; }
  void assign(int a, int i, const std::string& e, std::vector<PipeOpd> ops) {
    int t = (int)st.size();
    st.push_back({a, i, e, ops});
    int v = ++cur[{a, i}];
    producer[{a, i, v}] = t;
  }

  void emit(std::ostream& o, const PipeParams& pp) {
    int n = (int)st.size();
    auto vname = [](int a, int i, int v) {
      return "r" + std::to_string(a) + "_" + std::to_string(i) + "_" + std::to_string(v);
    };
    // operand sources
    struct Read {
      int use, earliest, arr, idx, ver;
    };
    std::vector<Read> reads;
    std::vector<std::vector<std::string>> opname(n);
    std::set<std::tuple<int, int, int>> must_store;
    for (int t = 0; t < n; t++) {
      for (auto& od : st[t].opds) {
        auto it = producer.find({od.arr, od.idx, od.ver});
        int prod = (it == producer.end()) ? -1 : it->second;
        if (prod >= 0 && t - prod <= pp.window) {
          opname[t].push_back(vname(od.arr, od.idx, od.ver));
        } else {
          opname[t].push_back("m" + std::to_string(reads.size()));
          reads.push_back({t, prod >= 0 ? prod + 1 : 0, od.arr, od.idx, od.ver});
          if (prod >= 0) must_store.insert({od.arr, od.idx, od.ver});
        }
      }
    }
    for (auto& kv : cur)  // final versions are the routine's output
      if (kv.second > 0) must_store.insert({kv.first.first, kv.first.second, kv.second});
    // last read used at each statement
    std::vector<int> last_read_at(n, -1), first_read_from(n + 1, (int)reads.size());
    for (int r = 0; r < (int)reads.size(); r++) last_read_at[reads[r].use] = r;
    for (int t = n - 1, r = (int)reads.size(); t >= 0; t--) {
      while (r > 0 && reads[r - 1].use >= t) r--;
      first_read_from[t] = r;
    }
    int issued = 0;
    std::map<std::pair<int, int>, int> ver_now;
    for (int t = 0; t < n; t++) {
      int target = std::min((int)reads.size(), first_read_from[t] + pp.dist);
      target = std::max(target, last_read_at[t] + 1);
      while (issued < target && reads[issued].earliest <= t) {
        auto& r = reads[issued];
        o << "  const double m" << issued << " = " << arr_name[r.arr] << "[" << r.idx << "];\n";
        issued++;
      }
      if (issued <= last_read_at[t]) die("pipeline: operand not loadable in time");
      auto& s = st[t];
      std::string e = s.expr;
      for (int k = (int)s.opds.size() - 1; k >= 0; k--) {
        std::string key = "$" + std::to_string(k);
        size_t p;
        while ((p = e.find(key)) != std::string::npos) e.replace(p, key.size(), opname[t][k]);
      }
      int v = ++ver_now[{s.arr, s.idx}];
      o << "  const double " << vname(s.arr, s.idx, v) << " = " << e << ";\n";
      if (must_store.count({s.arr, s.idx, v}))
        o << "  " << arr_name[s.arr] << "[" << s.idx << "] = " << vname(s.arr, s.idx, v) << ";\n";
      if ((t + 1) % pp.chunk == 0) o << "  __builtin_amdgcn_sched_barrier(0);\n";
    }
  }
};

void emit_pipelined(std::ostream& o, const Structure& S, const PipeParams& pp) {
  const int G = 0, X = 1;
  {  // KppDecomp (util/sutil.c:2-23), row-wise elimination in place
    Pipeline pl;
    pl.arr_name = {"G"};
    for (int k = 0; k < S.nvar; k++) {
      std::map<int, int> pos;
      for (int kk = S.lu_crow[k]; kk < S.lu_crow[k + 1]; kk++) pos[S.lu_icol[kk]] = kk;
      for (int kk = S.lu_crow[k]; kk < S.lu_diag[k]; kk++) {
        int j = S.lu_icol[kk];
        pl.assign(G, kk, "$0/$1", {pl.read(G, kk), pl.read(G, S.lu_diag[j])});
        for (int jj = S.lu_diag[j] + 1; jj < S.lu_crow[j + 1]; jj++) {
          int p = pos.at(S.lu_icol[jj]);
          pl.assign(G, p, "$0 - $1*$2", {pl.read(G, p), pl.read(G, kk), pl.read(G, jj)});
        }
      }
    }
    o << "// KppDecomp with memory pipelining (same operations, same order)\n";
    o << "template <class TG>\nstatic __device__ __forceinline__ void KppDecompPF(TG G) {\n";
    pl.emit(o, pp);
    o << "}\n";
  }
  {  // KppSolve (src/gen.c:1960-2033), one subtraction per statement
    Pipeline pl;
    pl.arr_name = {"G", "X"};
    for (int i = 0; i < S.nvar; i++)
      for (int j = S.lu_crow[i]; j < S.lu_diag[i]; j++)
        pl.assign(X, i, "$0 - $1*$2", {pl.read(X, i), pl.read(G, j), pl.read(X, S.lu_icol[j])});
    for (int i = S.nvar - 1; i >= 0; i--) {
      for (int j = S.lu_diag[i] + 1; j < S.lu_crow[i + 1]; j++)
        pl.assign(X, i, "$0 - $1*$2", {pl.read(X, i), pl.read(G, j), pl.read(X, S.lu_icol[j])});
      pl.assign(X, i, "$0/$1", {pl.read(X, i), pl.read(G, S.lu_diag[i])});
    }
    o << "// KppSolve with memory pipelining (same operations, same order)\n";
    o << "template <class TG, class TX>\nstatic __device__ __forceinline__ void KppSolvePF(TG G, TX X) {\n";
    pl.emit(o, pp);
    o << "}\n";
  }
}

void emit_coop(Model& md, std::ostream& o, const std::string& ns);

// The HIP header: everything is a template over an element accessor, so the
// same straight-line code runs on SoA global memory (p[i*stride]), LDS or
// register arrays.  All indices are compile-time constants.
void emit_hip(Model& md, const std::string& path) {
  Mech& m = *md.m;
  auto& S = md.S;
  std::string ns = "kppm_" + ident_ok(m.root);
  std::ofstream o(path);
  o << "// Generated by kppgen from the KPP mechanism '" << m.root << "'.\n"
    << "// HIP device functions of the model; arithmetic and association order are\n"
    << "// those of KPP's C output, so with FP contraction off the results are\n"
    << "// bitwise those of the C model.\n#pragma once\n#include <hip/hip_runtime.h>\n\n";
  if (g_lds_split > 0)
    o << "#ifndef KPPHIP_LDS_SPLIT_HELPERS\n#define KPPHIP_LDS_SPLIT_HELPERS\n"
      << "// Entry x of the LU with mirror slot R: in the LDS mirror Gs when R < NL,\n"
      << "// else in G.  R and NL are compile-time constants, so each access is one\n"
      << "// or the other.\n"
      << "template <int NL, int R, class TL, class TG>\n"
      << "static __device__ __forceinline__ typename TG::value_type kpp_lget(TL Gs, TG G, int x) {\n"
      << "  if constexpr (R < NL) return Gs[R]; else return G[x];\n}\n"
      << "template <int NL, int R, class TL, class TG>\n"
      << "static __device__ __forceinline__ void kpp_lput(TL Gs, TG G, int x,\n"
      << "                                                typename TG::value_type v) {\n"
      << "  if constexpr (R < NL) Gs[R] = v; else G[x] = v;\n}\n"
      << "template <int NL, int R, class TL, class TG>\n"
      << "static __device__ __forceinline__ void kpp_lcopy(TL Gs, TG G, int x) {\n"
      << "  if constexpr (R < NL) Gs[R] = G[x];\n}\n#endif\n\n";
  o << "struct " << ns << " {\n";
  o << "static constexpr int NVAR = " << S.nvar << ", NFIX = " << S.nfix << ", NSPEC = " << S.nspc
    << ", NREACT = " << S.nreact << ", NONZERO = " << S.nonzero
    << ", LU_NONZERO = " << S.lu_nonzero << ";\n";
  // LU entries KppDecompL/KppSolveL can mirror in LDS (slots 0..LDS_MAX-1)
  o << "static constexpr int LDS_MAX = " << std::min(g_lds_split, S.lu_nonzero) << ";\n";
  o << "static constexpr const char* ROOT = \"" << m.root << "\";\n";
  // rate dependence classes
  bool any_time = false;
  std::vector<int> rc_const, rc_sun, rc_state;
  for (auto& pr : md.rconst) {
    const int j = pr.first, cls = md.rate_class.at(j);
    if (md.f90) {
      auto d = md.f90->deps(md.f90_uses.at(j), md.f90->stmts.size());
      if (d.count("TIME")) any_time = true;
    } else if (mentions(pr.second, "TIME")) {
      any_time = true;
    }
    (cls == 2 ? rc_state : (cls == 1 ? rc_sun : rc_const)).push_back(j);
  }
  o << "static constexpr bool RATES_USE_TIME = " << (any_time ? "true" : "false") << ";\n";
  o << "static constexpr bool RATES_USE_STATE = " << (rc_state.empty() ? "false" : "true")
    << ";\n";
  o << "static constexpr int N_SUN_RATES = " << rc_sun.size() << ";\n";
  o << "static constexpr int N_STATE_RATES = " << rc_state.size() << ";\n";
  // #FUNCTION SPLIT: the F90 integrators call Fun_SPLIT (src/code.c:279-280,
  // 361-362: KPP_FUN_OR_FUN_SPLIT); the C integrators always call Fun
  o << "static constexpr bool FUN_SPLIT = " << (md.fun_split ? "true" : "false") << ";\n";
  if (md.fun_split) {
    o << "// Fun_SPLIT (src/gen.c:713-812): Vdot = P - D*V\n";
    o << "template <class TV, class TF, class TR, class TO>\nstatic __device__ __forceinline__ "
         "void FunSplit(TV V, TF F, TR RCT, TO Vdot) {\n";
    for (size_t q = 0; q < md.as_used.size(); q++)
      o << "  const double A" << md.as_used[q] << " = " << md.as_expr[q] << ";\n";
    auto locals = [](const std::string& e) {  // A[j] -> Aj
      std::string r;
      for (size_t p = 0; p < e.size(); p++) {
        if (e[p] == 'A' && p + 1 < e.size() && e[p + 1] == '[' &&
            (p == 0 || !(isalnum((unsigned char)e[p - 1]) || e[p - 1] == '_'))) {
          size_t q = e.find(']', p);
          r += "A" + e.substr(p + 2, q - p - 2);
          p = q;
        } else {
          r += e[p];
        }
      }
      return r;
    };
    for (int i = 0; i < S.nvar; i++)
      o << "  { const double P = " << locals(md.p_expr[i]) << "; const double D = "
        << md.d_expr[i] << "; Vdot[" << i << "] = P - D*V[" << i << "]; }\n";
    o << "}\n";
  }
  o << "#define KPPQ static __device__ __forceinline__\n" << kRateLawsBody << f90::kPowiHelpers
    << "#undef KPPQ\n";
  o << "// rate constants that depend on neither SUN, TIME nor the state (once per cell)\n";
  o << "template <class R> static __device__ __forceinline__ void rates_fixed(double TEMP, double "
       "CFACTOR, R RCT) {\n  (void)TEMP; (void)CFACTOR;\n";
  if (md.f90) md.f90->emit(o, md.preamble(rc_const));
  for (int j : rc_const) o << "  RCT[" << j << "] = (" << md.rate_c(j) << ");\n";
  o << "}\n// rate constants that depend on SUN or TIME (every Fun/Jac call)\n";
  o << "template <class R> static __device__ __forceinline__ void rates_sun(double TEMP, double "
       "CFACTOR, double SUN, double TIME, R RCT) {\n  (void)TEMP; (void)CFACTOR; (void)SUN; "
       "(void)TIME;\n";
  if (md.f90) md.f90->emit(o, md.preamble(rc_sun));
  for (int j : rc_sun) o << "  RCT[" << j << "] = (" << md.rate_c(j) << ");\n";
  o << "}\n// rate constants that read the concentrations (after every accepted step)\n";
  o << "template <class TV, class TF, class R> static __device__ __forceinline__ void "
       "rates_state(double TEMP, double CFACTOR, double SUN, double TIME, TV V, TF F, R RCT) {\n"
       "  (void)TEMP; (void)CFACTOR; (void)SUN; (void)TIME; (void)V; (void)F;\n";
  if (md.f90) md.f90->emit(o, md.preamble(rc_state));
  for (int j : rc_state) o << "  RCT[" << j << "] = (" << md.rate_c(j) << ");\n";
  o << "}\n";
  o << "template <class TV, class TF, class TR, class TO>\nstatic __device__ __forceinline__ void "
       "Fun(TV V, TF F, TR RCT, TO Vdot) {\n";
  {
    std::vector<std::string> chain(S.nvar);
    for (int i = 0; i < S.nvar; i++) {
      const std::string& e = md.vdot_expr[i];
      // A[j] -> Aj (locals)
      std::string r;
      for (size_t p = 0; p < e.size(); p++) {
        if (e[p] == 'A' && p + 1 < e.size() && e[p + 1] == '[' &&
            (p == 0 || !(isalnum((unsigned char)e[p - 1]) || e[p - 1] == '_'))) {
          size_t q = e.find(']', p);
          r += "A" + e.substr(p + 2, q - p - 2);
          p = q;
        } else {
          r += e[p];
        }
      }
      chain[i] = r;
    }
    // Reaction-major order.  Each Vdot[i] is a left-associated chain of terms
    // [-][c*]Aj in increasing j; evaluating the chains term by term right
    // after each Aj (a running value per species, stored when its last term
    // is in) performs the same operations on the same operands in the same
    // order, so the results are bitwise those of the chains, while only the
    // open running values, not all NREACT products, are live at once.
    struct Term { char op; std::string text; int j; };
    auto parse = [](const std::string& r, std::vector<Term>& out) {
      size_t p = 0;
      while (p < r.size()) {
        char op = 0;
        if (r[p] == '+' || r[p] == '-') op = r[p++];
        const size_t b = p;
        if (p < r.size() && (isdigit((unsigned char)r[p]) || r[p] == '.')) {
          while (p < r.size() && (isdigit((unsigned char)r[p]) || r[p] == '.')) p++;
          if (p < r.size() && (r[p] == 'e' || r[p] == 'E')) {
            p++;
            if (p < r.size() && (r[p] == '+' || r[p] == '-')) p++;
            while (p < r.size() && isdigit((unsigned char)r[p])) p++;
          }
          if (p >= r.size() || r[p] != '*') return false;
          p++;
        }
        if (p >= r.size() || r[p] != 'A') return false;
        p++;
        const size_t d = p;
        while (p < r.size() && isdigit((unsigned char)r[p])) p++;
        if (p == d) return false;
        Term t{op, r.substr(b, p - b), atoi(r.substr(d, p - d).c_str())};
        if (out.empty() && op == '-') t.text = "-" + t.text;  // leading unary minus
        if (out.empty()) t.op = 0;
        else if (op == 0) return false;
        if (!out.empty() && t.j <= out.back().j) return false;
        out.push_back(t);
      }
      return true;
    };
    std::vector<std::vector<Term>> terms(S.nvar);
    bool ok = true;
    for (int i = 0; i < S.nvar && ok; i++)
      if (chain[i].find('A') != std::string::npos) ok = parse(chain[i], terms[i]);
    if (ok) {
      // at[j]: (species, term index) pairs that use Aj
      std::map<int, std::vector<std::pair<int, int>>> at;
      for (int i = 0; i < S.nvar; i++)
        for (int t = 0; t < (int)terms[i].size(); t++) at[terms[i][t].j].push_back({i, t});
      for (int i = 0; i < S.nvar; i++)
        if (terms[i].empty()) o << "  Vdot[" << i << "] = " << chain[i] << ";\n";
      for (int i = 0; i < S.nvar; i++)
        if (terms[i].size() > 1) o << "  double f" << i << ";\n";
      for (size_t q = 0; q < md.a_used.size(); q++) {
        const int j = md.a_used[q];
        o << "  { const double A" << j << " = " << md.a_expr[q] << ";";
        for (auto& it : at[j]) {
          const int i = it.first;
          const Term& t = terms[i][it.second];
          const std::string v = "f" + std::to_string(i);
          if (t.op == 0)
            o << (terms[i].size() == 1 ? " Vdot[" + std::to_string(i) + "] = " + t.text + ";"
                                       : " " + v + " = " + t.text + ";");
          else
            o << " " << v << " = " << v << t.op << t.text << ";";
          if (terms[i].size() > 1 && it.second + 1 == (int)terms[i].size())
            o << " Vdot[" << i << "] = " << v << ";";
        }
        o << " }\n";
      }
    } else {
      for (size_t q = 0; q < md.a_used.size(); q++)
        o << "  const double A" << md.a_used[q] << " = " << md.a_expr[q] << ";\n";
      for (int i = 0; i < S.nvar; i++) o << "  Vdot[" << i << "] = " << chain[i] << ";\n";
    }
  }
  o << "}\n";
  // Jac_SP fused with the Ghimj build of ros_PrepareMatrix
  // (int/rosenbrock.c:686-704): G = -Jac, G[diag] += ghinv.  Returns the
  // first row (1-based) whose diagonal is exactly zero, the KppDecomp
  // singularity test of util/sutil.c:8, which reads the diagonal before that
  // row is eliminated.
  auto bsub = [&](const std::string& e) {
    std::string r;
    for (size_t p = 0; p < e.size(); p++) {
      if (e[p] == 'B' && p + 1 < e.size() && e[p + 1] == '[' &&
          (p == 0 || !(isalnum((unsigned char)e[p - 1]) || e[p - 1] == '_'))) {
        size_t q = e.find(']', p);
        r += "B" + e.substr(p + 2, q - p - 2);
        p = q;
      } else {
        r += e[p];
      }
    }
    return r;
  };
  // Large routines are emitted as a sequence of out-of-line parts of at
  // most g_part_ops operations (--part-ops): the compiler's time grows
  // faster than linearly with a function's length, and every value that
  // crosses a part boundary lives in the accessors' memory anyway.
  // kind: 0 void, 1 int with early return (first nonzero wins), 2 int whose
  // parts' results combine as the smallest nonzero value.
  auto routine = [&](const std::string& tparams, const std::string& name,
                     const std::string& params, const std::string& args, int kind,
                     const std::vector<std::pair<std::string, long>>& stmts,
                     const std::string& targs = "") {
    const char* rt = kind ? "int" : "void";
    std::vector<std::string> parts(1);
    long ops = 0;
    for (auto& st : stmts) {
      if (g_part_ops > 0 && ops >= g_part_ops && !parts.back().empty()) {
        parts.emplace_back();
        ops = 0;
      }
      parts.back() += st.first;
      ops += st.second;
    }
    const char* pre = kind == 2 ? "  int ising = 0;\n" : "";
    const char* post = kind == 1 ? "  return 0;\n" : (kind == 2 ? "  return ising;\n" : "");
    if (parts.size() == 1) {
      o << tparams << "\nstatic __device__ __forceinline__ " << rt << " " << name << "(" << params
        << ") {\n" << pre << parts[0] << post << "}\n";
      return;
    }
    for (size_t q = 0; q < parts.size(); q++)
      o << tparams << "\nstatic __device__ __noinline__ " << rt << " " << name << "_p" << q << "("
        << params << ") {\n" << pre << parts[q] << post << "}\n";
    o << tparams << "\nstatic __device__ __forceinline__ " << rt << " " << name << "(" << params
      << ") {\n";
    if (kind == 1) o << "  int r = 0;\n";
    if (kind == 2) o << "  int ising = 0, r = 0;\n";
    for (size_t q = 0; q < parts.size(); q++) {
      if (kind == 0) o << "  " << name << "_p" << q << targs << "(" << args << ");\n";
      if (kind == 1)
        o << "  if ((r = " << name << "_p" << q << targs << "(" << args << "))) return r;\n";
      if (kind == 2)
        o << "  r = " << name << "_p" << q << targs << "(" << args << ");\n"
          << "  if (r && (ising == 0 || r < ising)) ising = r;\n";
    }
    o << post << "}\n";
  };
  auto nops = [](const std::string& e) {
    long n = 0;
    for (char c : e) n += (c == '*' || c == '+' || c == '-' || c == '/');
    return n + 1;
  };
  // Jac_SP and JacGhimj, one Jacobian column at a time: the B terms of
  // column j (derivatives with respect to variable species j) feed only the
  // entries of column j, so only one column's B terms are live at once.
  // Every entry is the same expression, summed in the same order, as KPP's
  // Jac_SP (src/gen.c:1239-1284).
  std::vector<std::vector<int>> bcol(S.nvar), ncol(S.nvar);
  for (auto& kv : md.b_index) bcol[kv.first.second].push_back(kv.second);
  std::vector<int> lu_row(S.lu_nonzero);
  for (int i = 0; i < S.nvar; i++)
    for (int k = S.lu_crow[i]; k < S.lu_crow[i + 1]; k++) lu_row[k] = i;
  for (int k = 0; k < S.lu_nonzero; k++) ncol[S.lu_icol[k]].push_back(k);
  for (auto& v : bcol) std::sort(v.begin(), v.end());
  for (int ghimj = 0; ghimj < 2; ghimj++) {
    std::vector<std::pair<std::string, long>> st;
    for (int j = 0; j < S.nvar; j++) {
      std::string t = "  {\n";
      long n_ops = 0;
      for (int q : bcol[j]) {
        t += "  const double B" + std::to_string(q) + " = " + md.b_expr[q] + ";\n";
        n_ops += nops(md.b_expr[q]);
      }
      for (int n : ncol[j]) {
        std::string e = bsub(md.jvs_expr[n]);
        n_ops += nops(e);
        const int i = lu_row[n];
        if (!ghimj) {
          t += "  JVS[" + std::to_string(n) + "] = " + e + ";\n";
        } else if (n == S.lu_diag[i]) {
          // G = -Jac, G[diag] += ghinv (int/rosenbrock.c:686-704); the first
          // (smallest) row whose diagonal is exactly zero is reported, the
          // KppDecomp singularity test of util/sutil.c:8
          t += "  { const double g = -(" + e + ")+ghinv; G[" + std::to_string(n) +
               "] = g; if (g == 0.0 && (ising == 0 || ising > " + std::to_string(i + 1) +
               ")) ising = " + std::to_string(i + 1) + "; }\n";
        } else {
          t += "  G[" + std::to_string(n) + "] = -(" + e + ");\n";
        }
      }
      st.push_back({t + "  }\n", n_ops});
    }
    if (ghimj)
      routine("template <class TV, class TF, class TR, class TG>", "JacGhimj",
              "TV V, TF F, TR RCT, double ghinv, TG G", "V, F, RCT, ghinv, G", 2, st);
    else
      routine("template <class TV, class TF, class TR, class TJ>", "Jac_SP",
              "TV V, TF F, TR RCT, TJ JVS", "V, F, RCT, JVS", 0, st);
  }
  // KppDecomp unrolled: row-wise IKJ elimination of util/sutil.c:2-23 with
  // the work row W held in locals named after the row's LU entries (fill
  // closure guarantees every updated column is in the row's pattern).
  // l = W[j]/U_jj, W[c] -= l*U_jc is bitwise the reference's a=-W[j]/U_jj;
  // W[j]=-a; W[c]+=a*U_jc.
  // KppDecompChecked adds the singularity test of util/sutil.c:8 at the start
  // of each row and returns k+1 for the first row whose diagonal is zero.
  for (int checked = 0; checked < 2; checked++) {
    std::vector<std::pair<std::string, long>> st;
    for (int k = 0; k < S.nvar; k++) {
      std::string t;
      long n_ops = 0;
      if (checked) t += "  if (G[" + std::to_string(S.lu_diag[k]) + "] == 0.0) return " +
                        std::to_string(k + 1) + ";\n";
      std::map<int, int> pos;  // column -> position in row k
      for (int kk = S.lu_crow[k]; kk < S.lu_crow[k + 1]; kk++) pos[S.lu_icol[kk]] = kk;
      {
        // Row k held in locals: every entry the elimination touches is loaded
        // first and stored once at the end, so no store of row k sits between
        // the loads of the pivot rows it reads (earlier rows only) and they
        // can all be in flight together.  Per entry the operations and their
        // order are those of the IKJ loop: W[j] /= U_jj, then W[c] -= W[j]*U_jc.
        std::vector<int> touched;
        std::string ops;
        for (int kk = S.lu_crow[k]; kk < S.lu_diag[k]; kk++) {
          int j = S.lu_icol[kk];
          touched.push_back(kk);
          ops += "    w" + std::to_string(kk) + " = w" + std::to_string(kk) + "/G[" +
                 std::to_string(S.lu_diag[j]) + "];";
          n_ops++;
          for (int jj = S.lu_diag[j] + 1; jj < S.lu_crow[j + 1]; jj++) {
            auto it = pos.find(S.lu_icol[jj]);
            if (it == pos.end()) die("LU pattern not closed under elimination");
            const std::string w = "w" + std::to_string(it->second);
            touched.push_back(it->second);
            ops += " " + w + " = " + w + " - w" + std::to_string(kk) + "*G[" + std::to_string(jj) +
                   "];";
            n_ops += 2;
          }
          ops += "\n";
        }
        if (!touched.empty()) {
          std::sort(touched.begin(), touched.end());
          touched.erase(std::unique(touched.begin(), touched.end()), touched.end());
          t += "  {\n   ";
          for (int x : touched)
            t += " typename TG::value_type w" + std::to_string(x) + " = G[" + std::to_string(x) + "];";
          t += "\n" + ops + "   ";
          for (int x : touched) t += " G[" + std::to_string(x) + "] = w" + std::to_string(x) + ";";
          t += "\n  }\n";
        }
        st.push_back({t, n_ops});
      }
    }
    routine("template <class TG>", checked ? "KppDecompChecked" : "KppDecomp", "TG G", "G",
            checked ? 1 : 0, st);
  }
  // LDS-split variants (--lds-split N).  KppDecompL<NL>(G, Gs) and
  // KppSolveL<NL>(G, X, Gs) are KppDecomp and KppSolve with the factored
  // values of the NL entries KppDecomp re-reads most (pivot-row reads; slot 0
  // the most read) kept in Gs instead of G: the row-end store goes to Gs and
  // every later read of the entry (elimination of the rows below, the
  // triangular solves) comes from Gs.  G keeps the unfactored values of those
  // entries.  Same operations in the same order, so the results are bitwise
  // those of KppDecomp/KppSolve for every NL <= LDS_MAX.
  std::vector<int> lds_slot(S.lu_nonzero, -1);
  if (g_lds_split > 0) {
    std::vector<long> reads(S.lu_nonzero, 0);
    for (int k = 0; k < S.nvar; k++)
      for (int kk = S.lu_crow[k]; kk < S.lu_diag[k]; kk++) {
        const int j = S.lu_icol[kk];
        reads[S.lu_diag[j]]++;
        for (int jj = S.lu_diag[j] + 1; jj < S.lu_crow[j + 1]; jj++) reads[jj]++;
      }
    std::vector<int> by(S.lu_nonzero);
    for (int x = 0; x < S.lu_nonzero; x++) by[x] = x;
    std::stable_sort(by.begin(), by.end(), [&](int a, int b) { return reads[a] > reads[b]; });
    const int nl = std::min(g_lds_split, S.lu_nonzero);
    for (int r = 0; r < nl; r++) lds_slot[by[r]] = r;
  }
  auto lget = [&](const char* arr, int x) {
    if (lds_slot[x] < 0) return std::string(arr) + "[" + std::to_string(x) + "]";
    return "kpp_lget<NL, " + std::to_string(lds_slot[x]) + ">(Gs, " + arr + ", " +
           std::to_string(x) + ")";
  };
  if (g_lds_split > 0) {
    std::vector<std::pair<std::string, long>> st;
    for (int k = 0; k < S.nvar; k++) {
      std::string t;
      long n_ops = 0;
      std::map<int, int> pos;
      for (int kk = S.lu_crow[k]; kk < S.lu_crow[k + 1]; kk++) pos[S.lu_icol[kk]] = kk;
      std::vector<int> touched;
      std::string ops;
      for (int kk = S.lu_crow[k]; kk < S.lu_diag[k]; kk++) {
        const int j = S.lu_icol[kk];
        touched.push_back(kk);
        ops += "    w" + std::to_string(kk) + " = w" + std::to_string(kk) + "/" +
               lget("G", S.lu_diag[j]) + ";";
        n_ops++;
        for (int jj = S.lu_diag[j] + 1; jj < S.lu_crow[j + 1]; jj++) {
          auto it = pos.find(S.lu_icol[jj]);
          if (it == pos.end()) die("LU pattern not closed under elimination");
          const std::string w = "w" + std::to_string(it->second);
          touched.push_back(it->second);
          ops += " " + w + " = " + w + " - w" + std::to_string(kk) + "*" + lget("G", jj) + ";";
          n_ops += 2;
        }
        ops += "\n";
      }
      std::sort(touched.begin(), touched.end());
      touched.erase(std::unique(touched.begin(), touched.end()), touched.end());
      if (!touched.empty()) {
        t += "  {\n   ";
        for (int x : touched)
          t += " typename TG::value_type w" + std::to_string(x) + " = G[" + std::to_string(x) + "];";
        t += "\n" + ops + "   ";
        for (int x : touched) {
          const std::string w = "w" + std::to_string(x);
          if (lds_slot[x] < 0)
            t += " G[" + std::to_string(x) + "] = " + w + ";";
          else
            t += " kpp_lput<NL, " + std::to_string(lds_slot[x]) + ">(Gs, G, " +
                 std::to_string(x) + ", " + w + ");";
        }
        t += "\n  }\n";
      }
      // entries of row k the elimination leaves as they are: their mirror
      // slot gets the value from G
      for (int kk = S.lu_crow[k]; kk < S.lu_crow[k + 1]; kk++)
        if (lds_slot[kk] >= 0 && !std::binary_search(touched.begin(), touched.end(), kk))
          t += "  kpp_lcopy<NL, " + std::to_string(lds_slot[kk]) + ">(Gs, G, " +
               std::to_string(kk) + ");\n";
      st.push_back({t, n_ops});
    }
    routine("template <int NL, class TG, class TL>", "KppDecompL", "TG G, TL Gs", "G, Gs", 0, st,
            "<NL>");
  }
  if (g_emit_pipelined) emit_pipelined(o, S, g_pipe);
  // Residual of the unfactored matrix: R = B - G*X over the entries of the
  // pattern that can be nonzero (Jacobian entries and the diagonal; fill-in
  // entries of Ghimj are exactly zero).  Used by iterative refinement.
  {
    std::vector<std::pair<std::string, long>> st;
    for (int i = 0; i < S.nvar; i++) {
      std::string t = "  R[" + std::to_string(i) + "] = B[" + std::to_string(i) + "]";
      long n_ops = 0;
      for (int k = S.lu_crow[i]; k < S.lu_crow[i + 1]; k++)
        if (!md.jvs_zero[k] || k == S.lu_diag[i]) {
          t += " - G[" + std::to_string(k) + "]*X[" + std::to_string(S.lu_icol[k]) + "]";
          n_ops += 2;
        }
      st.push_back({t + ";\n", n_ops});
    }
    routine("template <class TG, class TX, class TB, class TR>", "Residual",
            "TG G, TX X, TB B, TR R", "G, X, B, R", 0, st);
  }
  // Iteration matrices from a stored Jacobian, for the integrators that keep
  // it between factorisations (SDIRK_PrepareMatrix, int/sdirk.c:905-960;
  // RK_Decomp, int/runge_kutta.c:1190-1246): G = -J with ghinv added to the
  // diagonal (the reference's two loops, entry by entry the same values),
  // and DiagFill for the imaginary part of RK_Decomp's complex matrix (v on
  // the diagonal, zero elsewhere).
  {
    std::vector<std::pair<std::string, long>> st, sd;
    for (int i = 0; i < S.nvar; i++) {
      std::string t, d;
      for (int k = S.lu_crow[i]; k < S.lu_crow[i + 1]; k++) {
        const std::string n = std::to_string(k);
        const bool diag = k == S.lu_diag[i];
        t += "  G[" + n + "] = -J[" + n + "]" + (diag ? "+ghinv" : "") + ";\n";
        d += "  G[" + n + "] = " + (diag ? "v" : "0.0") + ";\n";
      }
      st.push_back({t, (long)(S.lu_crow[i + 1] - S.lu_crow[i]) + 1});
      sd.push_back({d, 1});
    }
    routine("template <class TJ, class TG>", "JacToGhimj", "TJ J, double ghinv, TG G",
            "J, ghinv, G", 0, st);
    routine("template <class TG>", "DiagFill", "TG G, double v", "G, v", 0, sd);
  }
  // KppDecompCmplxR (util/sutil.c:28-58), the complex elimination of
  // RK_Decomp, unrolled like KppDecomp: row k's real and imaginary parts
  // held in locals (wr<x>, wi<x>) and stored once at the end.  Per entry the
  // reference's operations in its order: den = dR*dR + dI*dI,
  // ar = -(wr_j*dR + wi_j*dI)/den, ai = -(wi_j*dR - wr_j*dI)/den,
  // w_j = -a, then w_c = w_c + ar*U_c - ai*V_c (real) and
  // w_c + ar*V_c + ai*U_c (imaginary).  Returns k+1 for the first row whose
  // real or imaginary diagonal is exactly zero when its elimination starts.
  {
    std::vector<std::pair<std::string, long>> st;
    for (int k = 0; k < S.nvar; k++) {
      const std::string dk = std::to_string(S.lu_diag[k]), k1 = std::to_string(k + 1);
      std::string t = "  if (GR[" + dk + "] == 0.0) return " + k1 + ";\n  if (GI[" + dk +
                      "] == 0.0) return " + k1 + ";\n";
      long n_ops = 0;
      std::map<int, int> pos;
      for (int kk = S.lu_crow[k]; kk < S.lu_crow[k + 1]; kk++) pos[S.lu_icol[kk]] = kk;
      std::vector<int> touched;
      std::string ops;
      for (int kk = S.lu_crow[k]; kk < S.lu_diag[k]; kk++) {
        const int j = S.lu_icol[kk];
        const std::string d = std::to_string(S.lu_diag[j]), x = std::to_string(kk);
        touched.push_back(kk);
        ops += "    { const double den = GR[" + d + "]*GR[" + d + "] + GI[" + d + "]*GI[" + d +
               "]; const double ar = -(wr" + x + "*GR[" + d + "] + wi" + x + "*GI[" + d +
               "])/den; const double ai = -(wi" + x + "*GR[" + d + "] - wr" + x + "*GI[" + d +
               "])/den; wr" + x + " = -ar; wi" + x + " = -ai;";
        n_ops += 11;
        for (int jj = S.lu_diag[j] + 1; jj < S.lu_crow[j + 1]; jj++) {
          auto it = pos.find(S.lu_icol[jj]);
          if (it == pos.end()) die("LU pattern not closed under elimination");
          const std::string c = std::to_string(it->second), u = std::to_string(jj);
          touched.push_back(it->second);
          ops += " wr" + c + " = wr" + c + " + ar*GR[" + u + "] - ai*GI[" + u + "]; wi" + c +
                 " = wi" + c + " + ar*GI[" + u + "] + ai*GR[" + u + "];";
          n_ops += 8;
        }
        ops += " }\n";
      }
      if (!touched.empty()) {
        std::sort(touched.begin(), touched.end());
        touched.erase(std::unique(touched.begin(), touched.end()), touched.end());
        t += "  {\n   ";
        for (int x : touched)
          t += " double wr" + std::to_string(x) + " = GR[" + std::to_string(x) + "], wi" +
               std::to_string(x) + " = GI[" + std::to_string(x) + "];";
        t += "\n" + ops + "   ";
        for (int x : touched)
          t += " GR[" + std::to_string(x) + "] = wr" + std::to_string(x) + "; GI[" +
               std::to_string(x) + "] = wi" + std::to_string(x) + ";";
        t += "\n  }\n";
      }
      st.push_back({t, n_ops + 2});
    }
    routine("template <class TG>", "KppDecompCmplxR", "TG GR, TG GI", "GR, GI", 1, st);
  }
  // KppSolveCmplxR (util/sutil.c:63-87): forward substitution with the unit
  // lower factor, XR[i] -= (U*XR[c] - V*XI[c]) then XI[i] -= (U*XI[c] +
  // V*XR[c]) per entry; backward with the running sums and the division by
  // the complex diagonal written out as the reference does.
  {
    std::vector<std::pair<std::string, long>> st;
    for (int i = 0; i < S.nvar; i++) {
      const std::string is = std::to_string(i);
      std::string t;
      for (int j = S.lu_crow[i]; j < S.lu_diag[i]; j++) {
        const std::string u = std::to_string(j), c = std::to_string(S.lu_icol[j]);
        t += "  XR[" + is + "] = XR[" + is + "] - (GR[" + u + "]*XR[" + c + "] - GI[" + u +
             "]*XI[" + c + "]);\n  XI[" + is + "] = XI[" + is + "] - (GR[" + u + "]*XI[" + c +
             "] + GI[" + u + "]*XR[" + c + "]);\n";
      }
      if (!t.empty()) st.push_back({t, 8L * (S.lu_diag[i] - S.lu_crow[i])});
    }
    for (int i = S.nvar - 1; i >= 0; i--) {
      const std::string is = std::to_string(i), d = std::to_string(S.lu_diag[i]);
      std::string t = "  { double sr = XR[" + is + "], si = XI[" + is + "];";
      for (int j = S.lu_diag[i] + 1; j < S.lu_crow[i + 1]; j++) {
        const std::string u = std::to_string(j), c = std::to_string(S.lu_icol[j]);
        t += " sr = sr - (GR[" + u + "]*XR[" + c + "] - GI[" + u + "]*XI[" + c + "]); si = si - (GR[" +
             u + "]*XI[" + c + "] + GI[" + u + "]*XR[" + c + "]);";
      }
      t += " const double den = GR[" + d + "]*GR[" + d + "] + GI[" + d + "]*GI[" + d +
           "]; XR[" + is + "] = (sr*GR[" + d + "] + si*GI[" + d + "])/den; XI[" + is +
           "] = (si*GR[" + d + "] - sr*GI[" + d + "])/den; }\n";
      st.push_back({t, 8L * (S.lu_crow[i + 1] - S.lu_diag[i] - 1) + 11});
    }
    routine("template <class TG, class TX>", "KppSolveCmplxR", "TG GR, TG GI, TX XR, TX XI",
            "GR, GI, XR, XI", 0, st);
  }
  for (int split = 0; split < (g_lds_split > 0 ? 2 : 1); split++) {
    auto jvs = [&](int x) {
      return split ? lget("JVS", x) : "JVS[" + std::to_string(x) + "]";
    };
    std::vector<std::pair<std::string, long>> st;
    for (int i = 0; i < S.nvar; i++) {
      int b = S.lu_crow[i], e = S.lu_diag[i];
      if (b < e) {
        std::string t = "  X[" + std::to_string(i) + "] = X[" + std::to_string(i) + "]";
        for (int j = b; j < e; j++)
          t += "-" + jvs(j) + "*X[" + std::to_string(S.lu_icol[j]) + "]";
        st.push_back({t + ";\n", 2L * (e - b)});
      }
    }
    for (int i = S.nvar - 1; i >= 0; i--) {
      int b = S.lu_diag[i] + 1, e = S.lu_crow[i + 1];
      std::string t;
      if (b < e) {
        t = "  X[" + std::to_string(i) + "] = (X[" + std::to_string(i) + "]";
        for (int j = b; j < e; j++)
          t += "-" + jvs(j) + "*X[" + std::to_string(S.lu_icol[j]) + "]";
        t += ")/" + jvs(S.lu_diag[i]) + ";\n";
      } else {
        t = "  X[" + std::to_string(i) + "] = X[" + std::to_string(i) + "]/" +
            jvs(S.lu_diag[i]) + ";\n";
      }
      st.push_back({t, 2L * (e - b) + 1});
    }
    if (split)
      routine("template <int NL, class TG, class TX, class TL>", "KppSolveL",
              "TG JVS, TX X, TL Gs", "JVS, X, Gs", 0, st, "<NL>");
    else
      routine("template <class TG, class TX>", "KppSolve", "TG JVS, TX X", "JVS, X", 0, st);
  }
  // the discrete adjoint's model functions (Hessian, HessTR_Vec, Jac_SP_Vec,
  // JacTR_SP_Vec, KppSolveTR), the oracle's expressions
  o << "static constexpr bool ADJOINT = " << (md.adjoint ? "true" : "false") << ";\n";
  if (md.adjoint) {
    const size_t nh = md.hess_expr.size();
    o << "static constexpr int NHESS = " << std::max<size_t>(nh, 1) << ";\n";
    o << "template <class TV, class TF, class TR, class TH>\nstatic __device__ __forceinline__ "
         "void Hessian(TV V, TF F, TR RCT, TH HESS) {\n  double D2A["
      << std::max<size_t>(md.d2a_expr.size(), 1) << "];\n";
    for (size_t q = 0; q < md.d2a_expr.size(); q++)
      o << "  D2A[" << q << "] = " << md.d2a_expr[q] << ";\n";
    for (size_t q = 0; q < nh; q++) o << "  HESS[" << q << "] = " << md.hess_expr[q] << ";\n";
    if (nh == 0) o << "  HESS[0] = 0.0;\n";
    o << "}\n";
    o << "template <class TH, class T1, class T2, class TO>\nstatic __device__ __forceinline__ "
         "void HessTR_Vec(TH HESS, T1 U1, T2 U2, TO HTU) {\n";
    for (int i = 0; i < S.nvar; i++) o << "  HTU[" << i << "] = " << md.htu_expr[i] << ";\n";
    o << "}\n";
    o << "template <class TJ, class TU, class TO>\nstatic __device__ __forceinline__ "
         "void Jac_SP_Vec(TJ JVS, TU UV, TO JUV) {\n";
    for (int i = 0; i < S.nvar; i++) o << "  JUV[" << i << "] = " << md.juv_expr[i] << ";\n";
    o << "}\n";
    o << "template <class TJ, class TU, class TO>\nstatic __device__ __forceinline__ "
         "void JacTR_SP_Vec(TJ JVS, TU UV, TO JTUV) {\n";
    for (int i = 0; i < S.nvar; i++) o << "  JTUV[" << i << "] = " << md.jtuv_expr[i] << ";\n";
    o << "}\n";
    o << "template <class TJ, class TX>\nstatic __device__ __forceinline__ "
         "void KppSolveTR(TJ JVS, TX X, TX XX) {\n";
    for (auto& st : solve_tr_stmts(S)) o << st;
    o << "}\n";
  }
  o << "#undef ARR_abc\n#undef ARR_ab\n#undef ARR_ac\n#undef EP2\n#undef EP3\n#undef FALL\n";
  o << "};  // struct " << ns << "\n";
  o << "namespace " << ns << "_data {\n";
  emit_int_array(o, "static const", "LU_ICOL", S.lu_icol);
  emit_int_array(o, "static const", "LU_CROW", S.lu_crow);
  emit_int_array(o, "static const", "LU_DIAG", S.lu_diag);
  o << "static const char* const SPC_NAMES[" << S.nspc << "] = {";
  for (int k = 0; k < S.nspc; k++) o << (k % 8 ? " " : "\n  ") << "\"" << m.spc(m.code[k]).name << "\",";
  o << "};\n";
  // Initialize(): #INITVALUES times CFACTOR (src/gen.c:3252-3290)
  {
    auto iv = init_texts(md);
    o << "static const double CFACTOR = " << efmt(m.cfactor) << ";\n";
    o << "static const double INIT[" << S.nspc << "] = {";
    for (int k = 0; k < S.nspc; k++) o << (k % 4 ? " " : "\n  ") << "(" << iv[k] << ")*CFACTOR,";
    o << "};\n";
  }
  o << "}\n";
  emit_coop(md, o, ns);
}

// ---------------------------------------------------------------------------
// Cooperative programs (--coop): one wavefront integrates one cell, its 64
// lanes sharing the work of each model routine.  The routines become tables
// that a generic device interpreter walks step by step:
//  * products (A[j] of Fun, B[k] of Jac_SP): up to MAXF factors each, taken
//    left to right as in the emitted code (RCT[j], V[i], F[i] or a constant);
//    one product per lane at a time;
//  * sums and eliminations as chains: every value the model routines compute
//    as a left-associated chain (Vdot[i], JVS[n], an LU entry's updates and
//    its division, a row of the triangular solves) is one chain, whose steps
//    run consecutively on one lane with the running value in a register.
//    Same operands, same order as the emitted straight-line code, so the
//    results are bitwise the same.
// Chains are list-scheduled on 64 lanes in order of their longest path to
// the end; a step that reads another chain's result runs at least two steps
// after that chain's last step (the interpreter issues the loads of step
// t+1 before the store of step t).  Each table entry is one 64-bit
// descriptor per (step, lane); an all-ones descriptor is an idle lane.
// ---------------------------------------------------------------------------
bool g_coop = false;
bool g_coop_place = true;  // --coop-identity-layout: LU entry n at LDS slot n

namespace coop {

// An idle (lane, step) is all zeros; every real step has the VALID bit (52).
// Tables carry PAD idle steps past the last, so the device can request
// descriptors ahead without a bound check.
constexpr uint64_t IDLE = 0;
constexpr uint64_t VALID = 1ull << 52;
constexpr int PAD = 64;
// product factors: type in bits 30-31 (0 RCT, 1 V, 2 F, 3 constant), index below
constexpr uint32_t F_RCT = 0u << 30, F_V = 1u << 30, F_F = 2u << 30, F_C = 3u << 30;
constexpr uint32_t F_END = 0xFFFFFFFFu;

struct Term {
  double coef;
  int src;
};

// "[-][c*]X[n]" terms of a printed chain over array X ("A" or "B"); false if
// the text has another form
bool parse_chain(const std::string& e, const char* X, std::vector<Term>& out) {
  out.clear();
  if (e == "0.0" || e == "0") return true;
  size_t p = 0;
  const std::string ref = std::string(X) + "[";
  while (p < e.size()) {
    double sign = 1.0;
    if (e[p] == '+' || e[p] == '-') {
      if (e[p] == '-') sign = -1.0;
      p++;
    } else if (!out.empty()) {
      return false;
    }
    double c = 1.0;
    if (e.compare(p, ref.size(), ref) != 0) {
      char* end = nullptr;
      c = strtod(e.c_str() + p, &end);
      size_t q = end - e.c_str();
      if (q == p || q >= e.size() || e[q] != '*') return false;
      p = q + 1;
    }
    if (e.compare(p, ref.size(), ref) != 0) return false;
    p += ref.size();
    size_t q = e.find(']', p);
    if (q == std::string::npos) return false;
    out.push_back({sign * c, atoi(e.substr(p, q - p).c_str())});
    p = q + 1;
  }
  return true;
}

// a printed product "f*f*..." of RCT[j], V[i], F[i] and constants
bool parse_product(const std::string& e, std::vector<uint32_t>& fac, std::vector<double>& pool) {
  fac.clear();
  size_t p = 0;
  while (p <= e.size()) {
    size_t q = e.find('*', p);
    std::string f = e.substr(p, q == std::string::npos ? std::string::npos : q - p);
    auto idx = [&](const char* pre, uint32_t type) -> bool {
      size_t n = strlen(pre);
      if (f.compare(0, n, pre) != 0 || f.back() != ']') return false;
      fac.push_back(type | (uint32_t)atoi(f.substr(n, f.size() - n - 1).c_str()));
      return true;
    };
    if (!(idx("RCT[", F_RCT) || idx("V[", F_V) || idx("F[", F_F))) {
      char* end = nullptr;
      double v = strtod(f.c_str(), &end);
      if (f.empty() || end != f.c_str() + f.size() || f[0] == '-') return false;
      size_t k = 0;
      while (k < pool.size() && !(pool[k] == v)) k++;
      if (k == pool.size()) pool.push_back(v);
      fac.push_back(F_C | (uint32_t)k);
    }
    if (q == std::string::npos) break;
    p = q + 1;
  }
  return !fac.empty();
}

struct Chain {
  std::vector<uint64_t> steps;          // descriptors
  std::vector<uint64_t> steps2;         // paired programs: each step's second term (or IDLE)
  std::vector<std::vector<int>> deps;   // per step: chains whose results it reads
  int init_dep = -1;                    // chain producing the initial value (first step)
};

// An elimination chain as a paired program: steps 2k and 2k+1 become one
// step with two terms (the second IDLE for an odd chain's last), reading the
// operands of both.  Same operations in the same order; half the stores.
Chain pair_chain(const Chain& c) {
  Chain p;
  p.init_dep = c.init_dep;
  for (size_t k = 0; k < c.steps.size(); k += 2) {
    p.steps.push_back(c.steps[k]);
    std::vector<int> d = c.deps[k];
    if (k + 1 < c.steps.size()) {
      p.steps2.push_back(c.steps[k + 1]);
      d.insert(d.end(), c.deps[k + 1].begin(), c.deps[k + 1].end());
    } else {
      p.steps2.push_back(IDLE);
    }
    p.deps.push_back(d);
  }
  return p;
}

// List scheduling on 64 lanes; returns the table [T][64].
std::vector<uint64_t> schedule(const std::vector<Chain>& ch, int dist, int& T,
                               std::vector<uint64_t>* tab2 = nullptr) {
  const int n = (int)ch.size();
  std::vector<std::vector<int>> succ(n);
  for (int c = 0; c < n; c++) {
    std::set<int> d;
    for (auto& v : ch[c].deps) d.insert(v.begin(), v.end());
    if (ch[c].init_dep >= 0) d.insert(ch[c].init_dep);
    for (int x : d) succ[x].push_back(c);
  }
  // longest path from a chain's start to the end of the program
  std::vector<long> prio(n, -1);
  std::function<long(int)> pr = [&](int c) -> long {
    if (prio[c] >= 0) return prio[c];
    long m = 0;
    for (int s : succ[c]) m = std::max(m, pr(s));
    return prio[c] = (long)ch[c].steps.size() + dist + m;
  };
  for (int c = 0; c < n; c++) pr(c);
  std::vector<int> order(n);
  for (int c = 0; c < n; c++) order[c] = c;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return prio[a] > prio[b]; });
  std::vector<long> fin(n, -1), lane_free(64, 0);
  struct Place { int lane; long start; };
  std::vector<Place> at(n);
  long end = 0;
  for (int c : order) {
    if (ch[c].steps.empty()) die("coop: empty chain");
    long e = 0;
    if (ch[c].init_dep >= 0) {
      if (fin[ch[c].init_dep] < 0) die("coop: chain order violates a dependency");
      e = std::max(e, fin[ch[c].init_dep] + dist);
    }
    for (size_t k = 0; k < ch[c].deps.size(); k++)
      for (int d : ch[c].deps[k]) {
        if (fin[d] < 0) die("coop: chain order violates a dependency");
        e = std::max(e, fin[d] + dist - (long)k);
      }
    // the lane that can start it first; among those free before e, the latest free
    int best = -1;
    long bs = 0;
    for (int l = 0; l < 64; l++) {
      long s = std::max(lane_free[l], e);
      if (best < 0 || s < bs || (s == bs && lane_free[l] > lane_free[best])) {
        best = l;
        bs = s;
      }
    }
    at[c] = {best, bs};
    fin[c] = bs + (long)ch[c].steps.size() - 1;
    lane_free[best] = fin[c] + 1;
    end = std::max(end, fin[c] + 1);
  }
  T = (int)end;
  std::vector<uint64_t> tab((size_t)(T + PAD) * 64, IDLE);
  for (int c = 0; c < n; c++)
    for (size_t k = 0; k < ch[c].steps.size(); k++)
      tab[(size_t)(at[c].start + k) * 64 + at[c].lane] = ch[c].steps[k];
  if (tab2) {
    tab2->assign(tab.size(), IDLE);
    for (int c = 0; c < n; c++)
      for (size_t k = 0; k < ch[c].steps2.size(); k++)
        (*tab2)[(size_t)(at[c].start + k) * 64 + at[c].lane] = ch[c].steps2[k];
  }
  return tab;
}

// Scheduling descriptors (a program before encode_sums/encode_ops): 16-bit
// element indices, VALID bit 52.
// Sum chains (Vdot, JVS): constant-pool index of the coefficient bits 0-15
// (the coefficient itself is also tabled per step), source (A or B product)
// bits 16-31, target bits 32-47, FIRST 48, LAST 49, ZERO 50 (the value is
// 0.0), DIAG 51 (JVS: a diagonal entry, +ghinv).
uint64_t sum_desc(int coef, int src, int tgt, bool first, bool last, bool zero, bool diag) {
  return (uint64_t)coef | ((uint64_t)src << 16) | ((uint64_t)tgt << 32) |
         ((uint64_t)first << 48) | ((uint64_t)last << 49) | ((uint64_t)zero << 50) |
         ((uint64_t)diag << 51) | VALID;
}
// LU / solve steps: a bits 0-15, b bits 16-31, target bits 32-47, FIRST 48,
// LAST 49, DIV 50 (acc /= [a]; else acc -= [a]*[b])
uint64_t op_desc(int a, int b, int t, bool first, bool last, bool div) {
  return (uint64_t)a | ((uint64_t)b << 16) | ((uint64_t)t << 32) | ((uint64_t)first << 48) |
         ((uint64_t)last << 49) | ((uint64_t)div << 50) | VALID;
}

// ---- the device encoding of a scheduled program ----
// A cell's LDS (byte addresses; the kernel's dynamic LDS starts at 0):
//   G region [0, GZ), GZ = 8*LU_NONZERO: the LU entries, entry n at slot
//   GSLOT[n] (place_lds);
//   T region from GZ: T-relative 0 is a junk slot (JUNK), 8 holds +0.0
//   (ZERO), from 16 the products of Fun/Jac_SP (Fun's results after the A
//   products), the solve's X then Z vectors, the error-norm terms.
// Descriptor fields are 16-bit byte offsets from the region base the field
// belongs to, so the device adds no base (it is the LDS instruction's
// immediate offset).  Every step stores: a chain writes its running value to
// its own target (a partial value is never read: a target's original value
// is read only by the chain's own first step, other chains read it after the
// chain's last step); an idle step leaves the value unchanged (acc - 0*0, or
// acc + (-0.0)*0) and writes it back to the lane's previous target, or JUNK.
constexpr uint32_t T_JUNK = 0, T_ZERO = 8, T_DATA = 16;

// Sum programs, two terms of a chain per step: descriptor = product offset
// (T) | target offset << 16 (Fun: T, its results region at fo; Jac_SP: G) |
// second product offset << 32; bit 63 = FIRST.  Coefficients per step and
// term (two tables): the chain's, 1.0 for the single term of an empty chain
// (0.0 * 1.0), -0.0 for an absent term (its product is the +0.0 slot) and
// for idle steps.
void encode_sums(const std::vector<uint64_t>& tab, const std::vector<uint64_t>& tab2, bool jvs,
                 uint32_t gz, uint32_t fo, const std::vector<double>& pool,
                 const std::vector<int>& gslot, const std::vector<int>& pslot,
                 std::vector<uint64_t>& out, std::vector<double>& coef,
                 std::vector<double>& coef2) {
  const size_t steps = tab.size() / 64;
  out.assign(tab.size(), 0);
  coef.assign(tab.size(), 0.0);
  coef2.assign(tab.size(), 0.0);
  const uint32_t junk = jvs ? gz + T_JUNK : T_JUNK;
  for (int l = 0; l < 64; l++) {
    uint32_t prev = junk;
    for (size_t t = 0; t < steps; t++) {
      const uint64_t d = tab[t * 64 + l], d2 = tab2[t * 64 + l];
      uint32_t p = T_ZERO, p2 = T_ZERO, s = prev, first = 0;
      double c = -0.0, c2 = -0.0;
      if (d & VALID) {
        const bool zero = (d >> 50) & 1;
        const uint32_t src = (uint32_t)((d >> 16) & 0xFFFF), tgt = (uint32_t)((d >> 32) & 0xFFFF);
        p = zero ? T_ZERO : T_DATA + 8 * (uint32_t)pslot[src];
        c = zero ? 1.0 : pool[d & 0xFFFF];
        s = jvs ? 8 * (uint32_t)gslot[tgt] : fo + 8 * tgt;
        first = (d >> 48) & 1;
        prev = s;
        if (d2 & VALID) {
          if (((d2 >> 32) & 0xFFFF) != tgt || ((d2 >> 48) & 1) || ((d2 >> 50) & 1))
            die("coop: a pair spans two chains");
          p2 = T_DATA + 8 * (uint32_t)pslot[(d2 >> 16) & 0xFFFF];
          c2 = pool[d2 & 0xFFFF];
        }
      }
      out[t * 64 + l] = (uint64_t)p | ((uint64_t)s << 16) | ((uint64_t)p2 << 32) |
                        ((uint64_t)first << 63);
      coef[t * 64 + l] = c;
      coef2[t * 64 + l] = c2;
    }
  }
}

// Paired LU (every field in G) and solve (a in G; b, target and init in T,
// X at T_DATA, Z after it) programs, two words per (step, lane): lo = a1 |
// b1 << 16 | a2 << 32 | b2 << 48 (the step's two terms; an absent second
// term reads +0.0 twice), hi = target | init << 16, target bit 2 = the first
// term divides (acc / [a1]), bit 0 = the second does (acc / [a2]), bit 1 =
// some lane of the step divides (the same in all lanes); init = the first
// term's initial value when the step starts a chain, else JUNK (JUNK is the
// lowest address of the T region: solve inits are above it, LU inits below
// it).
void encode_ops(const std::vector<uint64_t>& tab, const std::vector<uint64_t>& tab2, bool solve,
                uint32_t gz, int nvar, const std::vector<int>& gslot, std::vector<uint64_t>& lo,
                std::vector<uint64_t>& hi) {
  const size_t steps = tab.size() / 64;
  lo.assign(tab.size(), 0);
  hi.assign(tab.size(), 0);
  const uint32_t junk = solve ? T_JUNK : gz + T_JUNK;
  const uint32_t zero_a = gz + T_ZERO, zero_b = solve ? T_ZERO : gz + T_ZERO;
  auto x_off = [&](uint32_t idx) { return solve ? T_DATA + 8 * idx : 8 * (uint32_t)gslot[idx]; };
  for (int l = 0; l < 64; l++) {
    uint32_t prev = junk;
    for (size_t t = 0; t < steps; t++) {
      const uint64_t d = tab[t * 64 + l], d2 = tab2[t * 64 + l];
      uint32_t a = zero_a, b = zero_b, a2 = zero_a, b2 = zero_b, s = prev, init = junk, flags = 0;
      if (d & VALID) {
        const uint32_t ai = (uint32_t)(d & 0xFFFF), bi = (uint32_t)((d >> 16) & 0xFFFF),
                       ti = (uint32_t)((d >> 32) & 0xFFFF);
        a = 8 * (uint32_t)gslot[ai];
        b = x_off(bi);
        s = x_off(ti);
        if ((d >> 48) & 1) {
          // backward solve chains (targets in Z) start from the forward result
          const uint32_t ii = (solve && ti >= (uint32_t)nvar) ? ti - (uint32_t)nvar : ti;
          init = x_off(ii);
        }
        if ((d >> 50) & 1) flags |= 4;
        if (d2 & VALID) {
          if (((d2 >> 32) & 0xFFFF) != ti || ((d2 >> 48) & 1)) die("coop: a pair spans two chains");
          a2 = 8 * (uint32_t)gslot[d2 & 0xFFFF];
          b2 = x_off((uint32_t)((d2 >> 16) & 0xFFFF));
          if ((d2 >> 50) & 1) flags |= 1;
        }
        prev = s;
      }
      lo[t * 64 + l] = (uint64_t)a | ((uint64_t)b << 16) | ((uint64_t)a2 << 32) |
                       ((uint64_t)b2 << 48);
      hi[t * 64 + l] = (uint64_t)(s | flags) | ((uint64_t)init << 16);
    }
  }
  for (size_t t = 0; t < steps; t++) {
    bool any = false;
    for (int l = 0; l < 64; l++) any |= (hi[t * 64 + l] & 5) != 0;
    if (any)
      for (int l = 0; l < 64; l++) hi[t * 64 + l] |= 2;
  }
}

// Placement of LDS arrays (the LU entries, the Fun and Jac_SP products).  A
// step's 64 lanes read up to 64 distinct elements per operand; elements whose
// 8-byte slots fall in the same bank pair are served one after the other.
// Each element gets a colour (its slot modulo BANKS8, the doubles per bank
// row) so that the elements one instruction touches spread over the colours:
// greedily, in the order of first use, the colour with the fewest elements
// of the element's instructions already on it, among the colours with free
// slots.  Returns slot[element], a permutation of 0..nnz-1.
constexpr int BANKS8 = 32;
std::vector<int> place_lds(int nnz, const std::vector<const std::vector<uint64_t>*>& progs,
                          const std::vector<int>& fields_of) {
  // instructions: (program, step, field) -> the distinct entries it touches
  std::vector<std::vector<int>> instr;
  std::vector<std::vector<int>> of_entry(nnz);
  std::vector<int> order;
  std::vector<char> seen(nnz, 0);
  for (size_t pi = 0; pi < progs.size(); pi++) {
    const auto& tab = *progs[pi];
    const size_t steps = tab.size() / 64;
    for (size_t t = 0; t < steps; t++)
      for (int f = 0; f < 3; f++) {
        if (!((fields_of[pi] >> f) & 1)) continue;
        std::set<int> es;
        for (int l = 0; l < 64; l++) {
          const uint64_t d = tab[t * 64 + l];
          if (d & VALID) es.insert((int)((d >> (16 * f)) & 0xFFFF));
        }
        if (es.size() < 2) {
          for (int e : es)
            if (!seen[e]) { seen[e] = 1; order.push_back(e); }
          continue;
        }
        const int id = (int)instr.size();
        instr.emplace_back(es.begin(), es.end());
        for (int e : es) {
          of_entry[e].push_back(id);
          if (!seen[e]) { seen[e] = 1; order.push_back(e); }
        }
      }
  }
  for (int e = 0; e < nnz; e++)
    if (!seen[e]) order.push_back(e);
  std::vector<int> cap(BANKS8), used(BANKS8, 0), colour(nnz, -1);
  for (int c = 0; c < BANKS8; c++) cap[c] = nnz / BANKS8 + (c < nnz % BANKS8 ? 1 : 0);
  std::vector<std::array<int, BANKS8>> cnt(instr.size());
  for (auto& a : cnt) a.fill(0);
  for (int e : order) {
    int best = -1;
    long bc = 0;
    for (int c = 0; c < BANKS8; c++) {
      if (used[c] >= cap[c]) continue;
      long cost = 0;
      for (int id : of_entry[e]) cost += cnt[id][c];
      if (best < 0 || cost < bc || (cost == bc && cap[c] - used[c] > cap[best] - used[best])) {
        best = c;
        bc = cost;
      }
    }
    colour[e] = best;
    for (int id : of_entry[e]) cnt[id][best]++;
    used[best]++;
  }
  std::vector<int> slot(nnz), next(BANKS8, 0);
  for (int e = 0; e < nnz; e++) slot[e] = colour[e] + BANKS8 * next[colour[e]]++;
  return slot;
}

void emit_table(std::ostream& o, const char* type, const std::string& name, size_t n,
                const std::function<std::string(size_t)>& item) {
  o << "static __device__ const " << type << " " << name << "[" << std::max<size_t>(n, 1)
    << "] = {";
  for (size_t i = 0; i < n; i++) o << (i % 8 ? " " : "\n  ") << item(i) << ",";
  if (n == 0) o << "0";
  o << "};\n";
}

}  // namespace coop

// The cooperative programs of a mechanism as `struct kppm_<root>_coop`
// (ENABLED = 0 when the mechanism's code has a form the tables cannot hold).
void emit_coop(Model& md, std::ostream& o, const std::string& ns) {
  using namespace coop;
  auto& S = md.S;
  const int nv = S.nvar, nnz = S.lu_nonzero;
  std::string why;
  std::vector<double> pool;
  std::vector<std::vector<uint32_t>> aprod, bprod;
  std::map<int, int> q_of;  // reaction -> A product index
  int maxf = 1;
  for (size_t q = 0; q < md.a_used.size() && why.empty(); q++) {
    std::vector<uint32_t> f;
    if (!parse_product(md.a_expr[q], f, pool)) why = "A product " + md.a_expr[q];
    aprod.push_back(f);
    q_of[md.a_used[q]] = (int)q;
    maxf = std::max(maxf, (int)f.size());
  }
  for (size_t k = 0; k < md.b_expr.size() && why.empty(); k++) {
    std::vector<uint32_t> f;
    if (!parse_product(md.b_expr[k], f, pool)) why = "B product " + md.b_expr[k];
    bprod.push_back(f);
    maxf = std::max(maxf, (int)f.size());
  }
  const int na = (int)aprod.size(), nb = (int)bprod.size();
  // T region (doubles): ZERO, JUNK, then the largest of the Fun products and
  // results, the Jac_SP products, the solve's X and Z
  const int tsize = 2 + std::max(std::max(na + nv, nb), 2 * nv);
  const uint32_t gz = 8u * (uint32_t)nnz, fo = T_DATA + 8u * (uint32_t)na;
  if (gz + T_JUNK > 0xFFFF || 8u * (uint32_t)tsize > 0xFFFF)
    why = "LDS byte offsets beyond 16 bits";
  auto pool_index = [&](double v) {
    size_t k = 0;
    while (k < pool.size() && !(pool[k] == v)) k++;
    if (k == pool.size()) pool.push_back(v);
    return (int)k;
  };
  std::vector<Chain> fun, jac, lu, sol;
  // Fun: Vdot[i] chains over A products
  for (int i = 0; i < nv && why.empty(); i++) {
    std::vector<Term> t;
    if (!parse_chain(md.vdot_expr[i], "A", t)) { why = "Vdot " + md.vdot_expr[i]; break; }
    Chain c;
    if (t.empty()) {
      c.steps.push_back(sum_desc(0, 0, i, true, true, true, false));
      c.deps.push_back({});
    }
    for (size_t k = 0; k < t.size(); k++) {
      auto it = q_of.find(t[k].src);
      if (it == q_of.end()) { why = "Vdot term"; break; }
      c.steps.push_back(sum_desc(pool_index(t[k].coef), it->second, i, k == 0, k + 1 == t.size(),
                                 false, false));
      c.deps.push_back({});
    }
    fun.push_back(c);
  }
  // Jac_SP fused with the Ghimj build: G[n] = -(JVS chain) (+ghinv on the diagonal)
  std::vector<char> is_diag(nnz, 0);
  for (int i = 0; i < nv; i++) is_diag[S.lu_diag[i]] = 1;
  for (int n = 0; n < nnz && why.empty(); n++) {
    std::vector<Term> t;
    if (!parse_chain(md.jvs_expr[n], "B", t)) { why = "JVS " + md.jvs_expr[n]; break; }
    Chain c;
    if (t.empty()) {
      c.steps.push_back(sum_desc(0, 0, n, true, true, true, is_diag[n]));
      c.deps.push_back({});
    }
    for (size_t k = 0; k < t.size(); k++) {
      c.steps.push_back(sum_desc(pool_index(t[k].coef), t[k].src, n, k == 0, k + 1 == t.size(),
                                 false, is_diag[n]));
      c.deps.push_back({});
    }
    jac.push_back(c);
  }
  // KppDecomp (util/sutil.c:2-23): entry x's chain = its updates by the pivot
  // rows in increasing order, then (L part) the division by the pivot
  if (why.empty()) {
    std::vector<int> chain_of(nnz, -1);
    std::vector<std::vector<uint64_t>> st(nnz);
    std::vector<std::vector<std::vector<int>>> dp(nnz);
    // dependencies are recorded as entries first, mapped to chains below
    for (int k = 0; k < nv; k++) {
      std::map<int, int> pos;
      for (int kk = S.lu_crow[k]; kk < S.lu_crow[k + 1]; kk++) pos[S.lu_icol[kk]] = kk;
      for (int kk = S.lu_crow[k]; kk < S.lu_diag[k]; kk++) {
        const int j = S.lu_icol[kk];
        st[kk].push_back(op_desc(S.lu_diag[j], 0, kk, false, false, true));
        dp[kk].push_back({S.lu_diag[j]});
        for (int jj = S.lu_diag[j] + 1; jj < S.lu_crow[j + 1]; jj++) {
          const int x = pos.at(S.lu_icol[jj]);
          st[x].push_back(op_desc(kk, jj, x, false, false, false));
          dp[x].push_back({kk, jj});
        }
      }
    }
    for (int x = 0; x < nnz; x++)
      if (!st[x].empty()) {
        chain_of[x] = (int)lu.size();
        lu.push_back({});
      }
    for (int x = 0; x < nnz; x++) {
      if (chain_of[x] < 0) continue;
      Chain& c = lu[chain_of[x]];
      for (size_t k = 0; k < st[x].size(); k++) {
        uint64_t d = st[x][k];
        if (k == 0) d |= 1ull << 48;
        if (k + 1 == st[x].size()) d |= 1ull << 49;
        c.steps.push_back(d);
        std::vector<int> dd;
        for (int e : dp[x][k])
          if (chain_of[e] >= 0) dd.push_back(chain_of[e]);
        c.deps.push_back(dd);
      }
    }
  }
  // KppSolve (the emitted forward and backward chains): forward row i in
  // place in X, backward row i from X[i] into Z[i] (b indices >= NVAR mean Z)
  if (why.empty()) {
    std::vector<int> fwd(nv, -1), bwd(nv, -1);
    for (int i = 0; i < nv; i++) {
      const int b = S.lu_crow[i], e = S.lu_diag[i];
      if (b == e) continue;
      Chain c;
      for (int j = b; j < e; j++) {
        const int col = S.lu_icol[j];
        c.steps.push_back(op_desc(j, col, i, j == b, j + 1 == e, false));
        c.deps.push_back(fwd[col] >= 0 ? std::vector<int>{fwd[col]} : std::vector<int>{});
      }
      fwd[i] = (int)sol.size();
      sol.push_back(c);
    }
    for (int i = nv - 1; i >= 0; i--) {
      const int b = S.lu_diag[i] + 1, e = S.lu_crow[i + 1];
      Chain c;
      c.init_dep = fwd[i];
      for (int j = b; j < e; j++) {
        const int col = S.lu_icol[j];
        c.steps.push_back(op_desc(j, nv + col, nv + i, j == b, false, false));
        c.deps.push_back({bwd[col]});
      }
      c.steps.push_back(op_desc(S.lu_diag[i], 0, nv + i, b == e, true, true));
      c.deps.push_back({});
      bwd[i] = (int)sol.size();
      sol.push_back(c);
    }
  }
  o << "// cooperative programs (one wavefront per cell; kppgen --coop)\n";
  if (!g_coop || !why.empty()) {
    if (g_coop)
      fprintf(stderr, "kppgen: no cooperative programs for %s: %s\n", md.m->root.c_str(),
              why.c_str());
    o << "struct " << ns << "_coop {\nstatic constexpr bool ENABLED = false;\n};\n";
    return;
  }
  int tf = 0, tj = 0, tl = 0, ts = 0;
  for (auto* v : {&fun, &jac, &lu, &sol})
    for (auto& c : *v) c = pair_chain(c);
  std::vector<uint64_t> tfun2, tjac2, tlu2, tsol2;
  auto tfun = schedule(fun, 2, tf, &tfun2), tjac = schedule(jac, 2, tj, &tjac2),
       tlu = schedule(lu, 2, tl, &tlu2), tsol = schedule(sol, 2, ts, &tsol2);
  auto hex = [](uint64_t v) {
    char b[32];
    snprintf(b, sizeof b, "0x%llxull", (unsigned long long)v);
    return std::string(b);
  };
  // product tables: factor-major, each factor row padded to a multiple of
  // 256 products with end codes (so four products per lane load unguarded)
  const int na_pad = (na + 255) / 256 * 256, nb_pad = (nb + 255) / 256 * 256;
  auto fac_item = [](const std::vector<std::vector<uint32_t>>& P, int npad) {
    return [&P, npad](size_t i) {
      const size_t f = i / npad, q = i % npad;
      const uint32_t v = (q < P.size() && f < P[q].size()) ? P[q][f] : F_END;
      return std::to_string(v) + "u";
    };
  };
  // LU entries' LDS slots: the LU program reads fields a, b and its
  // targets (bits 0-15, 16-31, 32-47), the solve field a, Jac_SP writes targets
  const std::vector<int> gslot =
      g_coop_place ? place_lds(nnz, {&tlu, &tlu2, &tsol, &tsol2, &tjac}, {7, 3, 1, 1, 4})
                   : [&] {
        std::vector<int> v(nnz);
        for (int n = 0; n < nnz; n++) v[n] = n;
        return v;
      }();
  {
    // the mean over instructions touching two or more entries of the
    // largest number of distinct entries on one bank pair
    auto load = [&](const std::vector<int>& sl) {
      const std::vector<const std::vector<uint64_t>*> progs = {&tlu, &tlu2, &tsol, &tsol2, &tjac};
      const int fields[] = {7, 3, 1, 1, 4};
      double tot = 0.0;
      long n = 0;
      for (size_t pi = 0; pi < progs.size(); pi++) {
        const auto& tab = *progs[pi];
        for (size_t t = 0; t < tab.size() / 64; t++)
          for (int f = 0; f < 3; f++) {
            if (!((fields[pi] >> f) & 1)) continue;
            std::set<int> es;
            for (int l = 0; l < 64; l++)
              if (tab[t * 64 + l] & VALID) es.insert((int)((tab[t * 64 + l] >> (16 * f)) & 0xFFFF));
            if (es.size() < 2) continue;
            int c[BANKS8] = {0}, mx = 0;
            for (int e : es) mx = std::max(mx, ++c[sl[e] % BANKS8]);
            tot += mx;
            n++;
          }
      }
      return n ? tot / n : 0.0;
    };
    std::vector<int> ident(nnz);
    for (int k = 0; k < nnz; k++) ident[k] = k;
    fprintf(stderr, "kppgen: %s coop LU bank-pair load per instruction: identity %.2f, placed %.2f\n",
            md.m->root.c_str(), load(ident), load(gslot));
  }
  // the products' slots (the sum programs read field b, bits 16-31): the
  // product tables are written in slot order, so the device's product loop
  // (lane-strided over slots) needs no map
  auto ident_of = [](int n) {
    std::vector<int> v(n);
    for (int k = 0; k < n; k++) v[k] = k;
    return v;
  };
  const std::vector<int> aslot = g_coop_place ? place_lds(na, {&tfun, &tfun2}, {2, 2}) : ident_of(na);
  const std::vector<int> bslot = g_coop_place ? place_lds(nb, {&tjac, &tjac2}, {2, 2}) : ident_of(nb);
  {
    std::vector<std::vector<uint32_t>> a2(na), b2(nb);
    for (int q = 0; q < na; q++) a2[aslot[q]] = aprod[q];
    for (int q = 0; q < nb; q++) b2[bslot[q]] = bprod[q];
    aprod.swap(a2);
    bprod.swap(b2);
  }
  std::vector<uint64_t> efun, ejac, elu, esol;
  std::vector<double> cfun, cjac, cfun2, cjac2;
  encode_sums(tfun, tfun2, false, gz, fo, pool, gslot, aslot, efun, cfun, cfun2);
  encode_sums(tjac, tjac2, true, gz, fo, pool, gslot, bslot, ejac, cjac, cjac2);
  std::vector<uint64_t> eluh, esolh;
  encode_ops(tlu, tlu2, false, gz, nv, gslot, elu, eluh);
  encode_ops(tsol, tsol2, true, gz, nv, gslot, esol, esolh);
  auto dbl = [](double v) {
    if (v == 0.0) return std::string(std::signbit(v) ? "-0.0" : "0.0");
    char b[40];
    snprintf(b, sizeof b, "%.17g", v);
    return std::string(b);
  };
  // the tables live in a namespace (device variables cannot be static data
  // members); the struct hands out their addresses
  o << "namespace " << ns << "_coop_data {\n";
  // product factors, factor-major: factor f of product q at [f*N + q]
  emit_table(o, "unsigned", "APROD", (size_t)maxf * na_pad, fac_item(aprod, na_pad));
  emit_table(o, "unsigned", "BPROD", (size_t)maxf * nb_pad, fac_item(bprod, nb_pad));
  emit_table(o, "double", "CONSTS", pool.size(), [&](size_t i) {
    char b[40];
    snprintf(b, sizeof b, "%.17g", pool[i]);
    return std::string(b);
  });
  emit_table(o, "unsigned long long", "FUN", efun.size(), [&](size_t i) { return hex(efun[i]); });
  emit_table(o, "unsigned long long", "JAC", ejac.size(), [&](size_t i) { return hex(ejac[i]); });
  emit_table(o, "double", "FUN_COEF", cfun.size(), [&](size_t i) { return dbl(cfun[i]); });
  emit_table(o, "double", "JAC_COEF", cjac.size(), [&](size_t i) { return dbl(cjac[i]); });
  emit_table(o, "double", "FUN_COEF2", cfun2.size(), [&](size_t i) { return dbl(cfun2[i]); });
  emit_table(o, "double", "JAC_COEF2", cjac2.size(), [&](size_t i) { return dbl(cjac2[i]); });
  emit_table(o, "unsigned long long", "LU", elu.size(), [&](size_t i) { return hex(elu[i]); });
  emit_table(o, "unsigned long long", "SOL", esol.size(), [&](size_t i) { return hex(esol[i]); });
  emit_table(o, "unsigned long long", "LUH", eluh.size(), [&](size_t i) { return hex(eluh[i]); });
  emit_table(o, "unsigned long long", "SOLH", esolh.size(), [&](size_t i) { return hex(esolh[i]); });
  emit_table(o, "unsigned", "DIAG", nv,
             [&](size_t i) { return std::to_string(gslot[S.lu_diag[i]]) + "u"; });
  emit_table(o, "unsigned", "GSLOT", nnz, [&](size_t i) { return std::to_string(gslot[i]) + "u"; });
  o << "}  // namespace " << ns << "_coop_data\n";
  o << "struct " << ns << "_coop {\n";
  o << "static constexpr bool ENABLED = true;\n";
  o << "static constexpr int NA = " << na << ", NB = " << nb << ", MAXF = " << maxf
    << ", NCONST = " << pool.size() << ", NA_PAD = " << na_pad << ", NB_PAD = " << nb_pad
    << ", PAD = " << PAD << ";\n";
  o << "// LDS byte offsets: G region size, Fun results (T-relative), T region doubles\n";
  o << "static constexpr unsigned GZ = " << gz << "u, FO = " << fo << "u, JSLOT = " << T_JUNK
    << "u, ZSLOT = " << T_ZERO << "u;\n";
  o << "static constexpr int TSIZE = " << tsize << ";\n";
  o << "static constexpr int FUN_STEPS = " << tf << ", JAC_STEPS = " << tj
    << ", LU_STEPS = " << tl << ", SOL_STEPS = " << ts << ";\n";
  const char* names[] = {"APROD", "BPROD", "CONSTS", "FUN", "JAC", "LU", "SOL", "DIAG",
                         "FUN_COEF", "JAC_COEF", "GSLOT", "LUH", "SOLH", "FUN_COEF2",
                         "JAC_COEF2"};
  const char* types[] = {"unsigned", "unsigned", "double", "unsigned long long",
                         "unsigned long long", "unsigned long long", "unsigned long long",
                         "unsigned", "double", "double", "unsigned", "unsigned long long",
                         "unsigned long long", "double", "double"};
  for (int k = 0; k < 15; k++)
    o << "static __device__ __forceinline__ const " << types[k] << "* " << names[k]
      << "() { return " << ns << "_coop_data::" << names[k] << "; }\n";
  o << "};\n";
  fprintf(stderr, "kppgen: %s coop steps: Fun %d, Jac %d, LU %d, solve %d (A %d, B %d)\n",
          md.m->root.c_str(), tf, tj, tl, ts, na, nb);
}

void emit_stats(Model& md, const std::string& path) {
  auto& S = md.S;
  std::ofstream o(path);
  long nused = (long)md.a_used.size();
  o << "{\n  \"root\": \"" << md.m->root << "\",\n  \"NVAR\": " << S.nvar << ",\n  \"NFIX\": "
    << S.nfix << ",\n  \"NSPEC\": " << S.nspc << ",\n  \"NREACT\": " << S.nreact
    << ",\n  \"NONZERO\": " << S.nonzero << ",\n  \"LU_NONZERO\": " << S.lu_nonzero
    << ",\n  \"NRATES_USED\": " << nused << ",\n  \"NB\": " << md.b_expr.size()
    << ",\n  \"flops\": {\"fun\": " << md.fun_ops.total() << ", \"jac\": " << md.jac_ops.total()
    << ", \"decomp\": " << md.lu_ops.total() << ", \"solve\": " << md.sol_ops.total()
    << ", \"decomp_div\": " << md.lu_ops.div << "},\n  \"species\": [";
  for (int k = 0; k < S.nspc; k++)
    o << (k ? ", " : "") << "\"" << md.m->spc(md.m->code[k]).name << "\"";
  o << "],\n  \"init\": [";
  auto iv = init_texts(md);
  for (int k = 0; k < S.nspc; k++) o << (k ? ", " : "") << "\"" << iv[k] << "\"";
  o << "],\n  \"cfactor\": " << efmt(md.m->cfactor);
  o << ",\n  \"lu_icol\": [";
  for (size_t i = 0; i < S.lu_icol.size(); i++) o << (i ? "," : "") << S.lu_icol[i];
  o << "],\n  \"lu_crow\": [";
  for (size_t i = 0; i < S.lu_crow.size(); i++) o << (i ? "," : "") << S.lu_crow[i];
  o << "],\n  \"lu_diag\": [";
  for (size_t i = 0; i < S.lu_diag.size(); i++) o << (i ? "," : "") << S.lu_diag[i];
  o << "],\n  \"rate_types\": [";
  for (int j = 0; j < S.nreact; j++) o << (j ? "," : "") << md.m->rates[j].type;
  o << "]\n}\n";
}

// KPP's C output files for host code that includes them
// (src/code_c.c:55-540): <root>_Parameters.h with the sizes and species
// indices of GenerateParamHeader (src/gen.c:2627-2700; C indices are
// 0-based, src/gen.c:4116-4121) as `#define NAME value` lines, and
// <root>_JacobianSP.c with the LU sparsity arrays of
// GenerateJacobianSparseData (src/gen.c:528-580, row-compressed LU pattern,
// 0-based).  Constants of features kppgen does not implement (families,
// #LOOKAT/#MONITOR/#CHECK lists, Hessians, stoichiometric matrices) are not
// written, so code that needs them fails to compile instead of using a
// wrong value.
void emit_kpp_c_files(Model& md, const std::string& base, const std::string& root) {
  auto& S = md.S;
  auto& m = *md.m;
  {
    std::ofstream o(base + "_Parameters.h");
    std::string guard = upper(root) + "_PARAMETERS_H_INCLUDED";
    o << "/* " << root << "_Parameters.h: sizes and species indices of mechanism " << root
      << ", in the layout of KPP's C output (written by kppgen) */\n";
    o << "#ifndef " << guard << "\n#define " << guard << "\n\n";
    auto def = [&](const std::string& n, long v, const char* what) {
      char b[160];
      snprintf(b, sizeof b, "#define %-20s %-10ld /* %s */\n", n.c_str(), v, what);
      o << b;
    };
    def("NSPEC", std::max(S.nspc, 1), "Number of chemical species");
    def("NVAR", std::max(S.nvar, 1), "Number of variable species");
    def("NFAM", std::max((int)m.families.size(), 1), "Number of prod/loss families");
    def("NFIX", std::max(S.nfix, 1), "Number of fixed species");
    def("NREACT", std::max(S.nreact, 1), "Number of reactions");
    def("NVARST", 0, "Index of the first variable species in C");
    def("NFIXST", S.nvar, "Index of the first fixed species in C");
    def("NONZERO", std::max(S.nonzero, 1), "Nonzero entries of the Jacobian");
    def("LU_NONZERO", std::max(S.lu_nonzero, 1), "Nonzero entries of its LU factors");
    def("CNVAR", S.nvar + 1, "NVAR+1: length of the row-compressed arrays");
    o << "\n/* Index declaration for variable species in C and VAR: VAR(ind_spc) = C(ind_spc) */\n";
    for (int k = 0; k < S.nvar; k++) def("ind_" + m.spc(m.code[k]).name, k, "variable");
    o << "\n/* Index declaration for fixed species in C: C(ind_spc) */\n";
    for (int k = S.nvar; k < S.nspc; k++) def("ind_" + m.spc(m.code[k]).name, k, "fixed");
    o << "\n/* Index declaration for fixed species in FIX: FIX(indf_spc) = C(NVAR+indf_spc) */\n";
    for (int k = S.nvar; k < S.nspc; k++)
      def("indf_" + m.spc(m.code[k]).name, k - S.nvar, "fixed, in FIX");
    o << "\n#endif\n";
  }
  {
    std::ofstream o(base + "_JacobianSP.c");
    o << "/* " << root << "_JacobianSP.c: row-compressed sparsity pattern of the LU factors of "
      << "the Jacobian of mechanism " << root << " (0-based), in the layout of KPP's C output "
      << "(written by kppgen) */\n\n";
    auto arr = [&](const char* name, const std::vector<int>& v, size_t n) {
      o << "  int  " << name << "[] = {\n     ";
      for (size_t i = 0; i < n; i++) {
        char b[16];
        snprintf(b, sizeof b, "%3d", v[i]);
        o << b << (i + 1 < n ? "," : "");
        if ((i + 1) % 12 == 0 && i + 1 < n) o << "\n     ";
      }
      o << " };\n\n";
    };
    arr("LU_IROW", S.lu_irow, (size_t)S.lu_nonzero);
    arr("LU_ICOL", S.lu_icol, (size_t)S.lu_nonzero);
    arr("LU_CROW", S.lu_crow, (size_t)S.nvar + 1);
    arr("LU_DIAG", S.lu_diag, (size_t)S.nvar + 1);
  }
}

}  // namespace

int main(int argc, char** argv) {
  std::string in, out_dir = ".", root, kpp_home, env;
  std::vector<std::string> f90_modules;
  bool do_reorder = true, rates_f90 = false, fun_split = false;
  for (int i = 1; i < argc; i++) {
    std::string a = argv[i];
    auto next = [&]() -> std::string {
      if (i + 1 >= argc) die("missing value for " + a);
      return argv[++i];
    };
    if (a == "-o") out_dir = next();
    else if (a == "--root") root = next();
    else if (a == "--kpp-home") kpp_home = next();
    else if (a == "--no-reorder") do_reorder = false;
    else if (a == "--rates-f90") rates_f90 = true;
    else if (a == "--no-pipelined") g_emit_pipelined = false;
    else if (a == "--pipelined") g_emit_pipelined = true;
    else if (a == "--coop") g_coop = true;
    else if (a == "--coop-identity-layout") g_coop_place = false;
    else if (a == "--part-ops") g_part_ops = atol(next().c_str());
    else if (a == "--lds-split") g_lds_split = atoi(next().c_str());
    else if (a == "--f90-module") f90_modules.push_back(next());
    else if (a == "--env") env = next();
    else if (a == "--pf-dist") g_pipe.dist = atoi(next().c_str());
    else if (a == "--pf-window") g_pipe.window = atoi(next().c_str());
    else if (a == "--pf-chunk") g_pipe.chunk = atoi(next().c_str());
    else if (a == "-h" || a == "--help") {
      printf("usage: kppgen [--kpp-home DIR] [--root NAME] [-o OUTDIR] [--rates-f90\n"
             "              [--f90-module FILE]... [--env 'F90 statements;...']]\n"
             "              [--part-ops N] [--lds-split N] [--pipelined]\n"
             "              mechanism.{def,kpp}\n");
      return 0;
    } else in = a;
  }
  if (in.empty()) die("no mechanism file given");
  if (root.empty()) {
    auto b = in.find_last_of('/');
    root = in.substr(b == std::string::npos ? 0 : b + 1);
    auto d = root.find_last_of('.');
    if (d != std::string::npos) root = root.substr(0, d);
  }
  if (kpp_home.empty() && getenv("KPP_HOME")) kpp_home = getenv("KPP_HOME");
  Scan sc;
  if (!kpp_home.empty()) sc.search.push_back(kpp_home + "/models");
  sc.file(in);

  Mech m;
  m.root = root;
  m.table.push_back({"???", DUMMY_SPC, ""});
  // Commands of src/scan.l:301-340.  Those that shape the model functions
  // are implemented or refused; those that only select KPP's other outputs
  // (drivers, integrator files, diagnostics, other languages) change nothing
  // the batched integrator computes and are skipped.  Anything else is an
  // error, so a mechanism never compiles into different equations silently.
  static const std::set<std::string> kNoEffect = {
      "", "INTEGRATOR", "HESSIAN", "DECLARE", "STOICMAT", "STOCHASTIC", "MEX", "DUMMYINDEX",
      "EQNTAGS", "ATOMS", "CHECK", "CHECKALL", "LOOKAT", "LOOKATALL", "MONITOR", "WRITE_ATM",
      "WRITE_SPC", "WRITE_MAT", "LANGUAGE", "INTFILE", "DRIVER", "UPPERCASEF90", "MINVERSION",
      "GRAPH", "ENDINLINE"};
  // parameter commands accepted only with the value the emitter implements
  auto param = [](const std::string& cmd, const std::string& body, const char* ok) {
    std::string v = upper(trim(body));
    while (!v.empty() && v.back() == ';') v = trim(v.substr(0, v.size() - 1));
    if (v != ok)
      die("#" + cmd + " " + v + " is not supported (only " + ok +
          "): it changes the model functions");
  };
  for (auto& s : sc.sections) {
    const std::string& c = s.cmd;
    if (c == "DEFVAR" || c == "DEFRAD") declare_species(m, s.body, VAR_SPC);
    else if (c == "DEFFIX") declare_species(m, s.body, FIX_SPC);
    else if (c == "SETVAR" || c == "SETRAD") set_type(m, s.body, VAR_SPC);
    else if (c == "SETFIX") set_type(m, s.body, FIX_SPC);
    else if (c == "INITVALUES") init_values(m, s.body);
    else if (c == "EQUATIONS")
      for (auto& e : split_semis(s.body)) parse_equation(m, e);
    else if (c == "REORDER") {
      std::string v = upper(trim(s.body));
      if (v.rfind("OFF", 0) == 0) do_reorder = false;
      else if (v.rfind("ON", 0) == 0) do_reorder = true;
      else die("#REORDER " + v + ": expected ON or OFF");
    } else if (c == "JACOBIAN") param(c, s.body, "SPARSE_LU_ROW");
    else if (c == "FUNCTION") {
      std::string v = upper(trim(s.body));
      while (!v.empty() && v.back() == ';') v = trim(v.substr(0, v.size() - 1));
      if (v == "SPLIT") fun_split = true;
      else if (v != "AGGREGATE") die("#FUNCTION " + v + ": expected AGGREGATE or SPLIT");
    }
    else if (c == "DOUBLE") param(c, s.body, "ON");
    else if (c == "FLUX") param(c, s.body, "OFF");
    else if (c == "AUTOREDUCE") param(c, s.body, "OFF");
    else if (c == "FAMILIES") declare_families(m, s.body);
    else if (!kNoEffect.count(c)) die("unknown or unsupported KPP command #" + c);
  }
  // inline code that enters the rate constants (src/gen.c Update_RCONST):
  // the F90 rate code is translated with --rates-f90; the C contexts are not
  // supported, and the other languages' code does not enter the C model
  for (auto& kv : sc.inlines) {
    const std::string& ctx = kv.first;
    if (!rates_f90 && (ctx == "C_RCONST" || ctx == "C_RATES"))
      die("#INLINE " + ctx + " is not supported: its C code would enter the rate constants");
    if (rates_f90 && ctx == "F90_RATES")
      die("#INLINE F90_RATES is not supported by the F90 rate translator");
  }
  if (m.rates.empty()) die("no equations");
  Model md;
  md.m = &m;
  md.fun_split = fun_split;
  md.S = reorder(m, do_reorder);
  md.build();
  // the adjoint's model functions, for mechanisms whose Hessian is modest
  md.adjoint = md.S.nvar <= 300;
  if (md.adjoint) md.build_adjoint();
  f90::Program prog;
  if (rates_f90) {
    // inputs of the rate code: the per-cell values of the C ABI
    prog.cx.inputs = {{"TEMP", "TEMP"}, {"CFACTOR", "CFACTOR"}, {"SUN", "SUN"}, {"TIME", "TIME"}};
    for (int k = 0; k < md.S.nspc; k++) prog.cx.species[upper(m.spc(m.code[k]).name)] = k;
    const int nv = md.S.nvar;
    prog.cx.conc = [nv](int k) {
      return k < nv ? "V[" + std::to_string(k) + "]" : "F[" + std::to_string(k - nv) + "]";
    };
    for (auto& f : f90_modules) prog.module(read_file(f));
    std::string envc = env;
    for (auto& ch : envc)
      if (ch == ';') ch = '\n';
    prog.code(envc);
    auto it = sc.inlines.find("F90_RCONST");
    if (it != sc.inlines.end()) prog.code(it->second);
    md.f90 = &prog;
  }
  md.classify();
  std::string base = out_dir + "/" + root;
  emit_c_model(md, base + "_model.h");
  emit_hip(md, base + "_hip.h");
  emit_stats(md, base + "_stats.json");
  emit_kpp_c_files(md, base, root);
  printf("kppgen: %s NVAR=%d NFIX=%d NREACT=%d NONZERO=%d LU_NONZERO=%d\n", root.c_str(),
         md.S.nvar, md.S.nfix, md.S.nreact, md.S.nonzero, md.S.lu_nonzero);
  return 0;
}
