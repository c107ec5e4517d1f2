// deck.hpp -- reader for the reference's input language (parser/Parser.y,
// parser/Lexer.l): nested "under Name { ... }" groups, "Key = value;"
// assignments, dotted and indexed names (A.B, A[3]), integer and double
// expressions with + - * / ^ and parentheses (integer operands give integer
// results, as in IntExpr), quoted strings, bare symbols (tokens such as
// True/False), // and /* */ comments, and an optional "end".
//
// The reference binds each key to a field of a registered object
// (ClassAssigner, parser/Assigner.h); here the result is a flat map from the
// full dotted name ("Diffusion.Temperature", "Lattice[0].PeriodicOA") to the
// value, which the host program reads by name.
#pragma once
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace a2c {

struct DeckValue {
  enum Kind { INT, DBL, STR, SYM } kind = INT;
  long long i = 0;
  double d = 0.0;
  std::string s;
  double num() const { return kind == INT ? (double)i : d; }
};

class Deck {
 public:
  static Deck parse_file(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open input file " + path);
    std::stringstream ss;
    ss << in.rdbuf();
    return parse(ss.str());
  }

  static Deck parse(const std::string& text) {
    Deck d;
    Parser p{text, 0, 1, &d};
    p.input({});
    return d;
  }

  bool has(const std::string& k) const { return v_.count(k) != 0; }
  double num(const std::string& k, double def) const {
    auto it = v_.find(k);
    return it == v_.end() ? def : it->second.num();
  }
  long long integer(const std::string& k, long long def) const {
    auto it = v_.find(k);
    return it == v_.end() ? def : (it->second.kind == DeckValue::INT ? it->second.i : (long long)it->second.d);
  }
  std::string str(const std::string& k, const std::string& def) const {
    auto it = v_.find(k);
    return it == v_.end() ? def : it->second.s;
  }
  const std::map<std::string, DeckValue>& values() const { return v_; }

 private:
  std::map<std::string, DeckValue> v_;

  struct Parser {
    const std::string& t;
    size_t p;
    int line;
    Deck* out;

    [[noreturn]] void fail(const std::string& m) {
      throw std::runtime_error("input line " + std::to_string(line) + ": " + m);
    }
    void skip() {
      for (;;) {
        while (p < t.size() && std::isspace((unsigned char)t[p])) {
          if (t[p] == '\n') ++line;
          ++p;
        }
        if (t.compare(p, 2, "//") == 0) {
          while (p < t.size() && t[p] != '\n') ++p;
        } else if (t.compare(p, 2, "/*") == 0) {
          size_t e = t.find("*/", p + 2);
          if (e == std::string::npos) fail("unterminated comment");
          for (size_t k = p; k < e; ++k) line += t[k] == '\n';
          p = e + 2;
        } else {
          return;
        }
      }
    }
    bool peek(char c) {
      skip();
      return p < t.size() && t[p] == c;
    }
    void expect(char c) {
      if (!peek(c)) fail(std::string("expected '") + c + "'");
      ++p;
    }
    std::string ident() {
      skip();
      if (p >= t.size() || !(std::isalpha((unsigned char)t[p]) || t[p] == '_')) fail("expected a name");
      size_t b = p;
      while (p < t.size() && (std::isalnum((unsigned char)t[p]) || t[p] == '_')) ++p;
      return t.substr(b, p - b);
    }
    std::string assignable() {
      std::string n = ident();
      for (;;) {
        if (peek('.')) {
          ++p;
          n += "." + ident();
        } else if (peek('[')) {
          ++p;
          DeckValue v = expr();
          if (v.kind != DeckValue::INT) fail("array index must be an integer");
          expect(']');
          n += "[" + std::to_string(v.i) + "]";
        } else {
          return n;
        }
      }
    }
    DeckValue number() {
      skip();
      size_t b = p;
      if (p < t.size() && (t[p] == '+' || t[p] == '-')) ++p;
      bool dbl = false;
      while (p < t.size() && std::isdigit((unsigned char)t[p])) ++p;
      if (p < t.size() && t[p] == '.') {
        dbl = true;
        ++p;
        while (p < t.size() && std::isdigit((unsigned char)t[p])) ++p;
      }
      if (p < t.size() && (t[p] == 'e' || t[p] == 'E' || t[p] == 'd' || t[p] == 'D')) {
        dbl = true;
        ++p;
        if (p < t.size() && (t[p] == '+' || t[p] == '-')) ++p;
        while (p < t.size() && std::isdigit((unsigned char)t[p])) ++p;
      }
      std::string s = t.substr(b, p - b);
      if (s.empty() || s == "+" || s == "-") fail("expected a number");
      for (char& ch : s)
        if (ch == 'd' || ch == 'D') ch = 'e';
      DeckValue v;
      if (dbl) { v.kind = DeckValue::DBL; v.d = std::atof(s.c_str()); }
      else { v.kind = DeckValue::INT; v.i = std::atoll(s.c_str()); }
      return v;
    }
    DeckValue factor() {
      if (peek('(')) {
        ++p;
        DeckValue v = expr();
        expect(')');
        return v;
      }
      return number();
    }
    static DeckValue arith(const DeckValue& a, const DeckValue& b, char op) {
      DeckValue r;
      if (a.kind == DeckValue::INT && b.kind == DeckValue::INT) {
        r.kind = DeckValue::INT;
        r.i = op == '+' ? a.i + b.i : op == '-' ? a.i - b.i : op == '*' ? a.i * b.i : (b.i ? a.i / b.i : 0);
      } else {
        r.kind = DeckValue::DBL;
        const double x = a.num(), y = b.num();
        r.d = op == '+' ? x + y : op == '-' ? x - y : op == '*' ? x * y : x / y;
      }
      return r;
    }
    // a ^ b: the reference's lexer passes '^' and Parser.y:29 gives it the
    // tightest binding, left-associative (%left '^'), but no rule of IntExpr
    // or DblExpr reduces it; here it is the power, integer when both sides
    // are integers and the exponent is not negative, double otherwise.
    static DeckValue power_of(const DeckValue& a, const DeckValue& b) {
      DeckValue r;
      if (a.kind == DeckValue::INT && b.kind == DeckValue::INT && b.i >= 0) {
        r.kind = DeckValue::INT;
        long long acc = 1, base = a.i;
        for (long long e = b.i; e; e >>= 1) {
          if (e & 1) acc *= base;
          base *= base;
        }
        r.i = acc;
      } else {
        r.kind = DeckValue::DBL;
        r.d = std::pow(a.num(), b.num());
      }
      return r;
    }
    DeckValue power() {
      DeckValue v = factor();
      while (peek('^')) {
        ++p;
        v = power_of(v, factor());
      }
      return v;
    }
    DeckValue term() {
      DeckValue v = power();
      while (peek('*') || peek('/')) {
        char op = t[p++];
        v = arith(v, power(), op);
      }
      return v;
    }
    DeckValue expr() {
      DeckValue v = term();
      for (;;) {
        skip();
        // a sign directly followed by a digit after a complete term is a binary operator
        if (p < t.size() && (t[p] == '+' || t[p] == '-')) {
          char op = t[p++];
          v = arith(v, term(), op);
        } else {
          return v;
        }
      }
    }
    DeckValue value() {
      skip();
      if (p < t.size() && t[p] == '"') {
        size_t e = t.find('"', p + 1);
        if (e == std::string::npos) fail("unterminated string");
        DeckValue v;
        v.kind = DeckValue::STR;
        v.s = t.substr(p + 1, e - p - 1);
        p = e + 1;
        return v;
      }
      if (p < t.size() && (std::isalpha((unsigned char)t[p]) || t[p] == '_')) {
        DeckValue v;
        v.kind = DeckValue::SYM;
        v.s = ident();
        return v;
      }
      return expr();
    }
    void input(const std::vector<std::string>& scope) {
      for (;;) {
        skip();
        if (p >= t.size()) return;
        if (t[p] == '}') return;
        size_t save = p;
        std::string w = ident();
        if (w == "end") { p = t.size(); return; }
        if (w == "under") {
          std::string name = assignable();
          expect('{');
          std::vector<std::string> inner = scope;
          inner.push_back(name);
          input(inner);
          expect('}');
          continue;
        }
        p = save;
        std::string key = assignable();
        expect('=');
        DeckValue v = value();
        expect(';');
        std::string full;
        for (const std::string& s : scope) full += s + ".";
        out->v_[full + key] = v;
      }
    }
  };
};

}  // namespace a2c
