// cexpr.cpp -- recursive-descent parser and evaluator of the material
// expressions declared in cexpr.h.
#include "cexpr.h"

#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace scuffx {

struct CExpr {
  enum Op { NUM, VAR_W, ADD, SUB, MUL, DIV, POW, NEG, FUN } op = NUM;
  std::complex<double> value;
  std::string fun;
  std::vector<CExprPtr> args;
};

namespace {
struct Parser {
  const std::string &s;
  const std::map<std::string, double> &consts;
  size_t p = 0;
  std::string err;
  Parser(const std::string &t, const std::map<std::string, double> &c) : s(t), consts(c) {}
  void skip() { while (p < s.size() && std::isspace((unsigned char)s[p])) p++; }
  CExprPtr make(CExpr::Op op, CExprPtr a = nullptr, CExprPtr b = nullptr)
  {
    auto e = std::make_shared<CExpr>();
    e->op = op;
    if (a) e->args.push_back(a);
    if (b) e->args.push_back(b);
    return e;
  }
  CExprPtr number(std::complex<double> v)
  {
    auto e = std::make_shared<CExpr>();
    e->value = v;
    return e;
  }
  // sum := term (('+'|'-') term)*
  CExprPtr sum()
  {
    CExprPtr a = term();
    for (;;) {
      if (!a) return nullptr;
      skip();
      if (p < s.size() && (s[p] == '+' || s[p] == '-')) {
        char c = s[p++];
        CExprPtr b = term();
        if (!b) return nullptr;
        a = make(c == '+' ? CExpr::ADD : CExpr::SUB, a, b);
      } else return a;
    }
  }
  // term := unary (('*'|'/') unary)*
  CExprPtr term()
  {
    CExprPtr a = unary();
    for (;;) {
      if (!a) return nullptr;
      skip();
      if (p < s.size() && (s[p] == '*' || s[p] == '/')) {
        char c = s[p++];
        CExprPtr b = unary();
        if (!b) return nullptr;
        a = make(c == '*' ? CExpr::MUL : CExpr::DIV, a, b);
      } else return a;
    }
  }
  // unary := ('+'|'-') unary | power
  CExprPtr unary()
  {
    skip();
    if (p < s.size() && (s[p] == '-' || s[p] == '+')) {
      char c = s[p++];
      CExprPtr a = unary();
      if (!a) return nullptr;
      return c == '-' ? make(CExpr::NEG, a) : a;
    }
    return power();
  }
  // power := primary ('^' unary)?   (right-associative: a^b^c = a^(b^c))
  CExprPtr power()
  {
    CExprPtr a = primary();
    if (!a) return nullptr;
    skip();
    if (p < s.size() && s[p] == '^') {
      p++;
      CExprPtr b = unary();
      if (!b) return nullptr;
      return make(CExpr::POW, a, b);
    }
    return a;
  }
  CExprPtr primary()
  {
    skip();
    if (p >= s.size()) { err = "unexpected end of expression"; return nullptr; }
    if (s[p] == '(') {
      p++;
      CExprPtr a = sum();
      if (!a) return nullptr;
      skip();
      if (p >= s.size() || s[p] != ')') { err = "missing )"; return nullptr; }
      p++;
      return a;
    }
    if (std::isdigit((unsigned char)s[p]) || s[p] == '.') {
      const char *b = s.c_str() + p;
      char *e = nullptr;
      double v = std::strtod(b, &e);
      if (e == b) { err = "bad number"; return nullptr; }
      p += (size_t)(e - b);
      return number(v);
    }
    if (std::isalpha((unsigned char)s[p]) || s[p] == '_') {
      size_t q = p;
      while (q < s.size() && (std::isalnum((unsigned char)s[q]) || s[q] == '_')) q++;
      std::string name = s.substr(p, q - p);
      p = q;
      skip();
      if (p < s.size() && s[p] == '(') {
        static const char *funs[] = {"sqrt", "exp", "log", "sin", "cos", "tan", "sinh", "cosh", "tanh", "abs"};
        bool known = false;
        for (const char *f : funs) known = known || name == f;
        if (!known) { err = "unknown function " + name; return nullptr; }
        p++;
        CExprPtr a = sum();
        if (!a) return nullptr;
        skip();
        if (p >= s.size() || s[p] != ')') { err = "missing ) after argument of " + name; return nullptr; }
        p++;
        auto e = std::make_shared<CExpr>();
        e->op = CExpr::FUN;
        e->fun = name;
        e->args.push_back(a);
        return e;
      }
      auto it = consts.find(name);
      if (it != consts.end()) return number(it->second);
      if (name == "w") return make(CExpr::VAR_W);
      if (name == "i" || name == "I") return number(std::complex<double>(0.0, 1.0));
      if (name == "pi") return number(M_PI);
      err = "unknown name " + name;
      return nullptr;
    }
    err = std::string("unexpected character '") + s[p] + "'";
    return nullptr;
  }
};
}  // namespace

std::string parse_cexpr(const std::string &text, const std::map<std::string, double> &constants, CExprPtr *out)
{
  Parser P(text, constants);
  CExprPtr e = P.sum();
  if (!e) return P.err.empty() ? "invalid expression" : P.err;
  P.skip();
  if (P.p != text.size()) return "junk after expression: " + text.substr(P.p);
  *out = e;
  return "";
}

std::complex<double> eval_cexpr(const CExpr &e, std::complex<double> w)
{
  switch (e.op) {
    case CExpr::NUM: return e.value;
    case CExpr::VAR_W: return w;
    case CExpr::ADD: return eval_cexpr(*e.args[0], w) + eval_cexpr(*e.args[1], w);
    case CExpr::SUB: return eval_cexpr(*e.args[0], w) - eval_cexpr(*e.args[1], w);
    case CExpr::MUL: return eval_cexpr(*e.args[0], w) * eval_cexpr(*e.args[1], w);
    case CExpr::DIV: return eval_cexpr(*e.args[0], w) / eval_cexpr(*e.args[1], w);
    case CExpr::NEG: return -eval_cexpr(*e.args[0], w);
    case CExpr::POW: {
      const std::complex<double> a = eval_cexpr(*e.args[0], w), b = eval_cexpr(*e.args[1], w);
      // integer powers by repeated products (exact for w^2 and the like)
      if (b.imag() == 0.0 && b.real() == std::round(b.real()) && std::fabs(b.real()) <= 64) {
        long n = (long)std::fabs(b.real());
        std::complex<double> r = 1.0, x = a;
        while (n) { if (n & 1) r *= x; x *= x; n >>= 1; }
        return b.real() < 0 ? 1.0 / r : r;
      }
      return std::pow(a, b);
    }
    case CExpr::FUN: {
      const std::complex<double> a = eval_cexpr(*e.args[0], w);
      const std::string &f = e.fun;
      if (f == "sqrt") return std::sqrt(a);
      if (f == "exp") return std::exp(a);
      if (f == "log") return std::log(a);
      if (f == "sin") return std::sin(a);
      if (f == "cos") return std::cos(a);
      if (f == "tan") return std::tan(a);
      if (f == "sinh") return std::sinh(a);
      if (f == "cosh") return std::cosh(a);
      if (f == "tanh") return std::tanh(a);
      return std::abs(a);
    }
  }
  return 0.0;
}

}  // namespace scuffx
