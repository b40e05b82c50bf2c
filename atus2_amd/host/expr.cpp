// expr.cpp -- recursive-descent parser to an expression tree (see expr.h).
#include "expr.h"

#include <cctype>
#include <cmath>
#include <cstdlib>
#include <functional>

namespace atus_host {

double sign_fn(double v) { return std::copysign(1.0, v); }
double heaviside_fn(double v) { return 0.5 * (1.0 + sign_fn(v)); }
double rect_fn(double x, double xl, double xr) { return heaviside_fn(x - xl) - heaviside_fn(x - xr); }

struct Expr::Node {
  enum Kind { NUM, VAR, UNARY, BINARY, TERNARY, CALL } kind = NUM;
  double value = 0.0;
  const double* var = nullptr;
  char op = 0;          // unary/binary operator; comparisons use codes below
  std::string fname;
  std::vector<std::shared_ptr<Node>> args;

  double eval() const {
    switch (kind) {
      case NUM: return value;
      case VAR: return *var;
      case UNARY: {
        double a = args[0]->eval();
        return op == '-' ? -a : a;
      }
      case TERNARY: return args[0]->eval() != 0.0 ? args[1]->eval() : args[2]->eval();
      case BINARY: {
        double a = args[0]->eval(), b = args[1]->eval();
        switch (op) {
          case '+': return a + b;
          case '-': return a - b;
          case '*': return a * b;
          case '/': return a / b;
          case '^': return std::pow(a, b);
          case '<': return a < b;
          case '>': return a > b;
          case 'l': return a <= b;
          case 'g': return a >= b;
          case '=': return a == b;
          case '!': return a != b;
          case '&': return (a != 0.0) && (b != 0.0);
          case '|': return (a != 0.0) || (b != 0.0);
        }
        return NAN;
      }
      case CALL: return call();
    }
    return NAN;
  }

  double call() const {
    const size_t n = args.size();
    auto a = [&](size_t i) { return args[i]->eval(); };
    const std::string& f = fname;
    if (n == 1) {
      double v = a(0);
      if (f == "sin") return std::sin(v);
      if (f == "cos") return std::cos(v);
      if (f == "tan") return std::tan(v);
      if (f == "asin") return std::asin(v);
      if (f == "acos") return std::acos(v);
      if (f == "atan") return std::atan(v);
      if (f == "sinh") return std::sinh(v);
      if (f == "cosh") return std::cosh(v);
      if (f == "tanh") return std::tanh(v);
      if (f == "asinh") return std::asinh(v);
      if (f == "acosh") return std::acosh(v);
      if (f == "atanh") return std::atanh(v);
      if (f == "exp") return std::exp(v);
      if (f == "log" || f == "ln") return std::log(v);
      if (f == "log2") return std::log2(v);
      if (f == "log10") return std::log10(v);
      if (f == "sqrt") return std::sqrt(v);
      if (f == "abs") return std::fabs(v);
      if (f == "rint") return std::rint(v);
      if (f == "sign") return sign_fn(v);
      if (f == "Heaviside") return heaviside_fn(v);
    }
    if (n == 3 && f == "rect") return rect_fn(a(0), a(1), a(2));
    if (f == "min" || f == "max" || f == "sum" || f == "avg") {
      double acc = a(0);
      for (size_t i = 1; i < n; ++i) {
        double v = a(i);
        if (f == "min") acc = v < acc ? v : acc;
        else if (f == "max") acc = v > acc ? v : acc;
        else acc += v;
      }
      return f == "avg" ? acc / (double)n : acc;
    }
    return NAN;
  }
};

namespace {

using NodeP = std::shared_ptr<Expr::Node>;

int arity_ok(const std::string& f, size_t n) {
  static const char* one[] = {"sin", "cos", "tan", "asin", "acos", "atan", "sinh", "cosh", "tanh", "asinh",
                              "acosh", "atanh", "exp", "log", "ln", "log2", "log10", "sqrt", "abs", "rint",
                              "sign", "Heaviside"};
  for (const char* o : one)
    if (f == o) return n == 1 ? 1 : -1;
  if (f == "rect") return n == 3 ? 1 : -1;
  if (f == "min" || f == "max" || f == "sum" || f == "avg") return n >= 1 ? 1 : -1;
  return 0;  // unknown function
}

class Parser {
 public:
  Parser(const std::string& s, const std::map<std::string, double>& c, const std::map<std::string, const double*>& v)
      : m_s(s), m_c(c), m_v(v) {}

  NodeP parse() {
    NodeP n = ternary();
    skip();
    if (m_p != m_s.size()) error("unexpected token");
    return n;
  }

 private:
  const std::string& m_s;
  const std::map<std::string, double>& m_c;
  const std::map<std::string, const double*>& m_v;
  size_t m_p = 0;

  [[noreturn]] void error(const std::string& what) {
    throw std::string("Error in expression \"" + m_s + "\" at position " + std::to_string(m_p) + ": " + what);
  }
  void skip() {
    while (m_p < m_s.size() && std::isspace((unsigned char)m_s[m_p])) ++m_p;
  }
  bool eat(const char* tok) {
    skip();
    size_t n = std::char_traits<char>::length(tok);
    if (m_s.compare(m_p, n, tok) == 0) {
      m_p += n;
      return true;
    }
    return false;
  }
  static NodeP bin(char op, NodeP a, NodeP b) {
    auto n = std::make_shared<Expr::Node>();
    n->kind = Expr::Node::BINARY;
    n->op = op;
    n->args = {a, b};
    return n;
  }
  NodeP ternary() {
    NodeP c = logic_or();
    if (eat("?")) {
      NodeP a = ternary();
      if (!eat(":")) error("expected ':'");
      NodeP b = ternary();
      auto n = std::make_shared<Expr::Node>();
      n->kind = Expr::Node::TERNARY;
      n->args = {c, a, b};
      return n;
    }
    return c;
  }
  NodeP logic_or() {
    NodeP a = logic_and();
    while (eat("||")) a = bin('|', a, logic_and());
    return a;
  }
  NodeP logic_and() {
    NodeP a = compare();
    while (eat("&&")) a = bin('&', a, compare());
    return a;
  }
  NodeP compare() {
    NodeP a = additive();
    for (;;) {
      if (eat("<=")) a = bin('l', a, additive());
      else if (eat(">=")) a = bin('g', a, additive());
      else if (eat("==")) a = bin('=', a, additive());
      else if (eat("!=")) a = bin('!', a, additive());
      else if (eat("<")) a = bin('<', a, additive());
      else if (eat(">")) a = bin('>', a, additive());
      else return a;
    }
  }
  NodeP additive() {
    NodeP a = multiplicative();
    for (;;) {
      if (eat("+")) a = bin('+', a, multiplicative());
      else if (eat("-")) a = bin('-', a, multiplicative());
      else return a;
    }
  }
  NodeP multiplicative() {
    NodeP a = unary();
    for (;;) {
      if (eat("*")) a = bin('*', a, unary());
      else if (eat("/")) a = bin('/', a, unary());
      else return a;
    }
  }
  NodeP unary() {
    if (eat("-")) {
      auto n = std::make_shared<Expr::Node>();
      n->kind = Expr::Node::UNARY;
      n->op = '-';
      n->args = {unary()};
      return n;
    }
    if (eat("+")) return unary();
    return power();
  }
  NodeP power() {
    NodeP base = primary();
    if (eat("^")) return bin('^', base, unary());  // right associative; exponent may carry a sign
    return base;
  }
  NodeP primary() {
    skip();
    if (m_p >= m_s.size()) error("unexpected end of expression");
    char ch = m_s[m_p];
    if (ch == '(') {
      ++m_p;
      NodeP n = ternary();
      if (!eat(")")) error("expected ')'");
      return n;
    }
    if (std::isdigit((unsigned char)ch) || ch == '.') {
      const char* b = m_s.c_str() + m_p;
      char* e = nullptr;
      double v = std::strtod(b, &e);
      if (e == b) error("bad number");
      m_p += (size_t)(e - b);
      auto n = std::make_shared<Expr::Node>();
      n->value = v;
      return n;
    }
    if (std::isalpha((unsigned char)ch) || ch == '_') {
      size_t b = m_p;
      while (m_p < m_s.size() && (std::isalnum((unsigned char)m_s[m_p]) || m_s[m_p] == '_')) ++m_p;
      std::string id = m_s.substr(b, m_p - b);
      skip();
      if (m_p < m_s.size() && m_s[m_p] == '(') {
        ++m_p;
        auto n = std::make_shared<Expr::Node>();
        n->kind = Expr::Node::CALL;
        n->fname = id;
        if (!eat(")")) {
          do n->args.push_back(ternary());
          while (eat(","));
          if (!eat(")")) error("expected ')' after arguments of " + id);
        }
        int ok = arity_ok(id, n->args.size());
        if (ok == 0) error("unknown function " + id);
        if (ok < 0) error("wrong number of arguments for " + id);
        return n;
      }
      auto vi = m_v.find(id);
      if (vi != m_v.end()) {
        auto n = std::make_shared<Expr::Node>();
        n->kind = Expr::Node::VAR;
        n->var = vi->second;
        return n;
      }
      auto ci = m_c.find(id);
      double v;
      if (ci != m_c.end()) v = ci->second;
      else if (id == "_pi") v = M_PI;
      else if (id == "_e") v = M_E;
      else error("unknown token " + id);
      auto n = std::make_shared<Expr::Node>();
      n->value = v;
      return n;
    }
    error(std::string("unexpected character '") + ch + "'");
  }
};

}  // namespace

Expr::Expr() = default;
Expr::~Expr() = default;
Expr::Expr(const Expr& o) = default;
Expr& Expr::operator=(const Expr& o) = default;

void Expr::DefineConst(const std::string& name, double value) {
  m_consts[name] = value;
  m_root.reset();
}
void Expr::ClearConst() {
  m_consts.clear();
  m_root.reset();
}
void Expr::DefineVar(const std::string& name, const double* ptr) {
  m_vars[name] = ptr;
  m_root.reset();
}
// parsing happens at the first Eval(), after the variables are bound (the
// reference binds them after SetExpr, e.g. rt_solver.cpp:88-100)
void Expr::SetExpr(const std::string& text) {
  m_text = text;
  m_root.reset();
}
double Expr::Eval() const {
  if (!m_root) {
    Parser p(m_text, m_consts, m_vars);
    const_cast<Expr*>(this)->m_root = p.parse();
  }
  return m_root->eval();
}

}  // namespace atus_host
