// Front end of the structured-product language: the lexer (src/dsl/
// lexer.rs:80-380), the recursive-descent parser (parser.rs:158-700), the
// compiler to the slot-resolved product (compiler.rs:40-330) and the device
// program builder (the register form of eval.rs:1542-1772). Error messages
// and their order follow the reference; spans are byte offsets as there.
#include "ofg_dsl_front.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>

namespace ofg_dsl {

namespace {

// ---------------------------------------------------------------------------
// Lexer
// ---------------------------------------------------------------------------

enum class Tk {
    Number, StringLit, True, False, Ident,
    Product, Notional, Maturity, Underlyings, State, Schedule, From, To, Let, If, Then, Else,
    Pay, Redeem, Set, Skip, And, Or, Not, Asset, Equity, Fx, Commodity, Rate,
    Monthly, Quarterly, SemiAnnual, Annual, Bool, Float,
    Indent, Dedent,
    LParen, RParen, Colon, Comma, Eq, EqEq, Ne, Lt, Le, Gt, Ge, Plus, Minus, Star, Slash,
};

struct Token {
    Tk kind;
    double num = 0.0;
    std::string text;
    size_t start = 0, end = 0;
};

// Rust's {:?} of an f64: shortest round-trip digits, ".0" on integers.
std::string debug_f64(double v) {
    if (std::isnan(v)) return "NaN";
    if (std::isinf(v)) return v > 0 ? "inf" : "-inf";
    char buf[64];
    for (int prec = 1; prec <= 17; ++prec) {
        std::snprintf(buf, sizeof buf, "%.*g", prec, v);
        if (std::strtod(buf, nullptr) == v) break;
    }
    std::string s(buf);
    if (s.find_first_of(".eEn") == std::string::npos) s += ".0";
    return s;
}

std::string debug_str(const std::string &s) {
    std::string o = "\"";
    for (char c : s) {
        if (c == '"' || c == '\\') o += '\\';
        if (c == '\n') { o += "\\n"; continue; }
        if (c == '\t') { o += "\\t"; continue; }
        o += c;
    }
    return o + "\"";
}

// TokenKind's Debug text (lexer.rs:14-78)
std::string debug_kind(const Token &t) {
    switch (t.kind) {
    case Tk::Number: return "Number(" + debug_f64(t.num) + ")";
    case Tk::StringLit: return "StringLit(" + debug_str(t.text) + ")";
    case Tk::Ident: return "Ident(" + debug_str(t.text) + ")";
    case Tk::True: return "True";
    case Tk::False: return "False";
    case Tk::Product: return "Product";
    case Tk::Notional: return "Notional";
    case Tk::Maturity: return "Maturity";
    case Tk::Underlyings: return "Underlyings";
    case Tk::State: return "State";
    case Tk::Schedule: return "Schedule";
    case Tk::From: return "From";
    case Tk::To: return "To";
    case Tk::Let: return "Let";
    case Tk::If: return "If";
    case Tk::Then: return "Then";
    case Tk::Else: return "Else";
    case Tk::Pay: return "Pay";
    case Tk::Redeem: return "Redeem";
    case Tk::Set: return "Set";
    case Tk::Skip: return "Skip";
    case Tk::And: return "And";
    case Tk::Or: return "Or";
    case Tk::Not: return "Not";
    case Tk::Asset: return "Asset";
    case Tk::Equity: return "Equity";
    case Tk::Fx: return "Fx";
    case Tk::Commodity: return "Commodity";
    case Tk::Rate: return "Rate";
    case Tk::Monthly: return "Monthly";
    case Tk::Quarterly: return "Quarterly";
    case Tk::SemiAnnual: return "SemiAnnual";
    case Tk::Annual: return "Annual";
    case Tk::Bool: return "Bool";
    case Tk::Float: return "Float";
    case Tk::Indent: return "Indent";
    case Tk::Dedent: return "Dedent";
    case Tk::LParen: return "LParen";
    case Tk::RParen: return "RParen";
    case Tk::Colon: return "Colon";
    case Tk::Comma: return "Comma";
    case Tk::Eq: return "Eq";
    case Tk::EqEq: return "EqEq";
    case Tk::Ne: return "Ne";
    case Tk::Lt: return "Lt";
    case Tk::Le: return "Le";
    case Tk::Gt: return "Gt";
    case Tk::Ge: return "Ge";
    case Tk::Plus: return "Plus";
    case Tk::Minus: return "Minus";
    case Tk::Star: return "Star";
    case Tk::Slash: return "Slash";
    }
    return "?";
}

std::string kind_name(Tk k) {
    Token t;
    t.kind = k;
    return debug_kind(t);
}

[[noreturn]] void lex_error(const std::string &m, size_t s, size_t e) {
    throw DslError(DslError::Lex,
                   "lex error at " + std::to_string(s) + "-" + std::to_string(e) + ": " + m);
}

[[noreturn]] void parse_error(const std::string &m, size_t s, size_t e) {
    throw DslError(DslError::Parse,
                   "parse error at " + std::to_string(s) + "-" + std::to_string(e) + ": " + m);
}

[[noreturn]] void compile_error(const std::string &m, bool has_span, size_t s, size_t e) {
    if (has_span)
        throw DslError(DslError::Compile, "compile error at " + std::to_string(s) + "-" +
                                              std::to_string(e) + ": " + m);
    throw DslError(DslError::Compile, "compile error: " + m);
}

[[noreturn]] void eval_error(const std::string &m) {
    throw DslError(DslError::Eval, "eval error: " + m);
}

bool is_digit(char c) { return c >= '0' && c <= '9'; }
bool is_alpha(char c) { return (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z'); }

void lex_content(const std::string &content, size_t offset, std::vector<Token> &out) {
    static const std::map<std::string, Tk> keywords = {
        {"product", Tk::Product}, {"notional", Tk::Notional}, {"maturity", Tk::Maturity},
        {"underlyings", Tk::Underlyings}, {"state", Tk::State}, {"schedule", Tk::Schedule},
        {"from", Tk::From}, {"to", Tk::To}, {"let", Tk::Let}, {"if", Tk::If},
        {"then", Tk::Then}, {"else", Tk::Else}, {"pay", Tk::Pay}, {"redeem", Tk::Redeem},
        {"set", Tk::Set}, {"skip", Tk::Skip}, {"and", Tk::And}, {"or", Tk::Or},
        {"not", Tk::Not}, {"asset", Tk::Asset}, {"equity", Tk::Equity}, {"fx", Tk::Fx},
        {"commodity", Tk::Commodity}, {"rate", Tk::Rate}, {"true", Tk::True},
        {"false", Tk::False}, {"monthly", Tk::Monthly}, {"quarterly", Tk::Quarterly},
        {"semi_annual", Tk::SemiAnnual}, {"annual", Tk::Annual}, {"bool", Tk::Bool},
        {"float", Tk::Float}};
    const size_t n = content.size();
    size_t pos = 0;
    while (pos < n) {
        while (pos < n && content[pos] == ' ') ++pos;
        if (pos >= n) break;
        if (pos + 1 < n && content[pos] == '/' && content[pos + 1] == '/') break;
        const size_t start = offset + pos;
        const char ch = content[pos];
        Token t;
        t.start = start;
        if (ch == '"') {   // lex_string (lexer.rs:283-310)
            size_t p = pos + 1;
            std::string s;
            bool closed = false;
            while (p < n) {
                if (content[p] == '"') {
                    closed = true;
                    break;
                }
                if (content[p] == '\\' && p + 1 < n) {
                    ++p;
                    switch (content[p]) {
                    case 'n': s += '\n'; break;
                    case 't': s += '\t'; break;
                    case '"': s += '"'; break;
                    case '\\': s += '\\'; break;
                    default: s += content[p];
                    }
                } else {
                    s += content[p];
                }
                ++p;
            }
            if (!closed) lex_error("unterminated string literal", offset + pos, offset + p);
            t.kind = Tk::StringLit;
            t.text = s;
            t.end = offset + p + 1;
            out.push_back(t);
            pos = p + 1;
            continue;
        }
        if (is_digit(ch) || (ch == '.' && pos + 1 < n && is_digit(content[pos + 1]))) {
            // lex_number (lexer.rs:312-364)
            size_t p = pos;
            std::string num;
            while (p < n && (is_digit(content[p]) || content[p] == '_')) {
                if (content[p] != '_') num += content[p];
                ++p;
            }
            if (p < n && content[p] == '.') {
                num += '.';
                ++p;
                while (p < n && (is_digit(content[p]) || content[p] == '_')) {
                    if (content[p] != '_') num += content[p];
                    ++p;
                }
            }
            if (p < n && (content[p] == 'e' || content[p] == 'E')) {
                num += 'e';
                ++p;
                if (p < n && (content[p] == '+' || content[p] == '-')) num += content[p++];
                while (p < n && is_digit(content[p])) num += content[p++];
            }
            // Rust's f64 parser: digits with optional fraction and exponent;
            // "1." and ".5" parse, an exponent without digits does not.
            bool ok = !num.empty() && num != ".";
            const size_t e = num.find('e');
            if (ok && e != std::string::npos) {
                size_t d = e + 1;
                if (d < num.size() && (num[d] == '+' || num[d] == '-')) ++d;
                ok = d < num.size();
                if (ok && e == 0) ok = false;
            }
            char *endp = nullptr;
            const double v = ok ? std::strtod(num.c_str(), &endp) : 0.0;
            if (!ok || !endp || *endp != '\0')
                lex_error("invalid number literal '" + num + "'", offset + pos, offset + p);
            t.kind = Tk::Number;
            t.num = v;
            t.end = offset + p;
            out.push_back(t);
            pos = p;
            continue;
        }
        if (is_alpha(ch) || ch == '_') {
            size_t p = pos;
            while (p < n && (is_alpha(content[p]) || is_digit(content[p]) || content[p] == '_')) ++p;
            const std::string word = content.substr(pos, p - pos);
            auto it = keywords.find(word);
            if (it != keywords.end()) {
                t.kind = it->second;
            } else {
                t.kind = Tk::Ident;
                t.text = word;
            }
            t.end = offset + p;
            out.push_back(t);
            pos = p;
            continue;
        }
        if (pos + 1 < n) {
            const std::string two = content.substr(pos, 2);
            Tk k;
            bool hit = true;
            if (two == "==") k = Tk::EqEq;
            else if (two == "!=") k = Tk::Ne;
            else if (two == "<=") k = Tk::Le;
            else if (two == ">=") k = Tk::Ge;
            else hit = false;
            if (hit) {
                t.kind = k;
                t.end = offset + pos + 2;
                out.push_back(t);
                pos += 2;
                continue;
            }
        }
        Tk k;
        switch (ch) {
        case '(': k = Tk::LParen; break;
        case ')': k = Tk::RParen; break;
        case ':': k = Tk::Colon; break;
        case ',': k = Tk::Comma; break;
        case '=': k = Tk::Eq; break;
        case '<': k = Tk::Lt; break;
        case '>': k = Tk::Gt; break;
        case '+': k = Tk::Plus; break;
        case '-': k = Tk::Minus; break;
        case '*': k = Tk::Star; break;
        case '/': k = Tk::Slash; break;
        default: {
            // the whole UTF-8 character, as Rust's chars().next()
            size_t len = 1;
            const unsigned char u = (unsigned char)ch;
            if (u >= 0xF0) len = 4;
            else if (u >= 0xE0) len = 3;
            else if (u >= 0xC0) len = 2;
            lex_error("unexpected character '" + content.substr(pos, len) + "'", start, start + 1);
        }
        }
        t.kind = k;
        t.end = offset + pos + 1;
        out.push_back(t);
        ++pos;
    }
}

// tokenize (lexer.rs:80-150): INDENT / DEDENT from leading spaces.
std::vector<Token> tokenize(const std::string &src) {
    std::vector<Token> toks;
    std::vector<size_t> stack{0};
    size_t global = 0;
    size_t i = 0;
    while (true) {
        const size_t nl = src.find('\n', i);
        const std::string line = src.substr(i, nl == std::string::npos ? std::string::npos : nl - i);
        const size_t line_start = global;
        std::string clean = line;
        while (!clean.empty() && clean.back() == '\r') clean.pop_back();
        size_t indent = 0;
        while (indent < clean.size() && clean[indent] == ' ') ++indent;
        std::string content = clean.substr(indent);
        // trim_end: whitespace (spaces, tabs, ...)
        while (!content.empty() && (content.back() == ' ' || content.back() == '\t' ||
                                    content.back() == '\r' || content.back() == '\f' ||
                                    content.back() == '\v'))
            content.pop_back();
        if (!(content.empty() || content.rfind("//", 0) == 0)) {
            const size_t cur = stack.back();
            if (indent > cur) {
                stack.push_back(indent);
                Token t;
                t.kind = Tk::Indent;
                t.start = line_start;
                t.end = line_start + indent;
                toks.push_back(t);
            } else if (indent < cur) {
                while (stack.back() > indent) {
                    stack.pop_back();
                    Token t;
                    t.kind = Tk::Dedent;
                    t.start = line_start;
                    t.end = line_start + indent;
                    toks.push_back(t);
                }
                if (stack.back() != indent)
                    lex_error("inconsistent indentation: expected " + std::to_string(stack.back()) +
                                  " spaces, got " + std::to_string(indent),
                              line_start, line_start + indent);
            }
            lex_content(content, line_start + indent, toks);
        }
        global += line.size() + 1;
        if (nl == std::string::npos) break;
        i = nl + 1;
    }
    while (stack.size() > 1) {
        stack.pop_back();
        Token t;
        t.kind = Tk::Dedent;
        t.start = t.end = global;
        toks.push_back(t);
    }
    return toks;
}

// ---------------------------------------------------------------------------
// Syntax tree (ast.rs) and parser
// ---------------------------------------------------------------------------

struct AExpr {
    enum Kind { Num, BoolLit, Ident, Bin, Un, Call } kind;
    double num = 0.0;
    bool b = false;
    std::string name;
    BinOp bop = BinOp::Add;
    UnOp uop = UnOp::Neg;
    std::vector<std::unique_ptr<AExpr>> args;
    size_t start = 0, end = 0;
};

struct AStmt {
    enum Kind { Let, If, Pay, Redeem, Set, Skip } kind;
    std::string name;
    std::unique_ptr<AExpr> expr;
    std::vector<AStmt> then_body, else_body;
    size_t start = 0, end = 0;
};

struct ASchedule {
    double period;   // ScheduleFreq::period (ast.rs:54-62)
    double start, end;
    std::vector<AStmt> body;
};

struct AUnderlying {
    std::string name;
    int64_t index;
    UnderlyingType type;
    size_t start, end;
};

struct AState {
    std::string name, type_name;
    std::unique_ptr<AExpr> init;
    size_t start, end;
};

struct AItem {
    enum Kind { Notional, Maturity, Underlyings, State, Schedule } kind;
    double value = 0.0;
    size_t start = 0, end = 0;
    std::vector<AUnderlying> und;
    std::vector<AState> state;
    ASchedule sched;
};

struct AProduct {
    std::string name;
    std::vector<AItem> body;
};

struct Parser {
    std::vector<Token> t;
    size_t pos = 0;

    const Token *peek() const { return pos < t.size() ? &t[pos] : nullptr; }
    bool at(Tk k) const { return pos < t.size() && t[pos].kind == k; }
    bool at_block_end() const { return pos >= t.size() || t[pos].kind == Tk::Dedent; }
    void advance() {
        if (pos < t.size()) ++pos;
    }
    size_t eof_pos() const { return t.empty() ? 0 : t.back().end; }
    size_t cur_start() const { return pos < t.size() ? t[pos].start : eof_pos(); }
    size_t cur_end() const { return pos < t.size() ? t[pos].end : eof_pos(); }
    std::string peek_debug() const {
        return pos < t.size() ? "Some(" + debug_kind(t[pos]) + ")" : "None";
    }

    const Token &expect(Tk k) {
        if (pos < t.size() && t[pos].kind == k) return t[pos++];
        if (pos < t.size())
            parse_error("expected " + kind_name(k) + ", got " + debug_kind(t[pos]), t[pos].start,
                        t[pos].end);
        parse_error("expected " + kind_name(k) + ", got end of input", eof_pos(), eof_pos());
    }
    const Token &expect_number() {
        if (pos < t.size() && t[pos].kind == Tk::Number) return t[pos++];
        if (pos < t.size())
            parse_error("expected number, got " + debug_kind(t[pos]), t[pos].start, t[pos].end);
        parse_error("expected number, got end of input", eof_pos(), eof_pos());
    }
    const Token &expect_string() {
        if (pos < t.size() && t[pos].kind == Tk::StringLit) return t[pos++];
        if (pos < t.size())
            parse_error("expected string, got " + debug_kind(t[pos]), t[pos].start, t[pos].end);
        parse_error("expected string, got end of input", eof_pos(), eof_pos());
    }
    // expect_ident (parser.rs:95-131): notional / maturity are accepted too.
    std::string expect_ident(size_t &s, size_t &e) {
        if (pos < t.size()) {
            const Token &k = t[pos];
            s = k.start;
            e = k.end;
            if (k.kind == Tk::Ident) {
                ++pos;
                return k.text;
            }
            if (k.kind == Tk::Notional) {
                ++pos;
                return "notional";
            }
            if (k.kind == Tk::Maturity) {
                ++pos;
                return "maturity";
            }
            parse_error("expected identifier, got " + debug_kind(k), k.start, k.end);
        }
        parse_error("expected identifier, got end of input", eof_pos(), eof_pos());
    }

    // ---- expressions (parser.rs:500-700) ----
    std::unique_ptr<AExpr> bin(BinOp op, std::unique_ptr<AExpr> l, std::unique_ptr<AExpr> r) {
        auto e = std::make_unique<AExpr>();
        e->kind = AExpr::Bin;
        e->bop = op;
        e->start = l->start;
        e->end = r->end;
        e->args.push_back(std::move(l));
        e->args.push_back(std::move(r));
        return e;
    }
    std::unique_ptr<AExpr> expr() { return or_expr(); }
    std::unique_ptr<AExpr> or_expr() {
        auto l = and_expr();
        while (at(Tk::Or)) {
            advance();
            l = bin(BinOp::Or, std::move(l), and_expr());
        }
        return l;
    }
    std::unique_ptr<AExpr> and_expr() {
        auto l = not_expr();
        while (at(Tk::And)) {
            advance();
            l = bin(BinOp::And, std::move(l), not_expr());
        }
        return l;
    }
    std::unique_ptr<AExpr> not_expr() {
        if (at(Tk::Not)) {
            const size_t s = cur_start();
            advance();
            auto o = not_expr();
            auto e = std::make_unique<AExpr>();
            e->kind = AExpr::Un;
            e->uop = UnOp::Not;
            e->start = s;
            e->end = o->end;
            e->args.push_back(std::move(o));
            return e;
        }
        return comparison();
    }
    std::unique_ptr<AExpr> comparison() {
        auto l = additive();
        BinOp op;
        if (at(Tk::EqEq)) op = BinOp::Eq;
        else if (at(Tk::Ne)) op = BinOp::Ne;
        else if (at(Tk::Lt)) op = BinOp::Lt;
        else if (at(Tk::Le)) op = BinOp::Le;
        else if (at(Tk::Gt)) op = BinOp::Gt;
        else if (at(Tk::Ge)) op = BinOp::Ge;
        else return l;
        advance();
        return bin(op, std::move(l), additive());
    }
    std::unique_ptr<AExpr> additive() {
        auto l = multiplicative();
        while (at(Tk::Plus) || at(Tk::Minus)) {
            const BinOp op = at(Tk::Plus) ? BinOp::Add : BinOp::Sub;
            advance();
            l = bin(op, std::move(l), multiplicative());
        }
        return l;
    }
    std::unique_ptr<AExpr> multiplicative() {
        auto l = unary();
        while (at(Tk::Star) || at(Tk::Slash)) {
            const BinOp op = at(Tk::Star) ? BinOp::Mul : BinOp::Div;
            advance();
            l = bin(op, std::move(l), unary());
        }
        return l;
    }
    std::unique_ptr<AExpr> unary() {
        if (at(Tk::Minus)) {
            const size_t s = cur_start();
            advance();
            auto o = primary();
            auto e = std::make_unique<AExpr>();
            e->kind = AExpr::Un;
            e->uop = UnOp::Neg;
            e->start = s;
            e->end = o->end;
            e->args.push_back(std::move(o));
            return e;
        }
        return primary();
    }
    std::unique_ptr<AExpr> primary() {
        const size_t s = cur_start(), e0 = cur_end();
        auto e = std::make_unique<AExpr>();
        e->start = s;
        e->end = e0;
        const Token *k = peek();
        if (k && k->kind == Tk::Number) {
            e->kind = AExpr::Num;
            e->num = k->num;
            advance();
            return e;
        }
        if (k && (k->kind == Tk::True || k->kind == Tk::False)) {
            e->kind = AExpr::BoolLit;
            e->b = k->kind == Tk::True;
            advance();
            return e;
        }
        if (k && k->kind == Tk::LParen) {
            advance();
            auto inner = expr();
            expect(Tk::RParen);
            return inner;
        }
        if (k && k->kind == Tk::Ident) {
            const std::string name = k->text;
            advance();
            if (at(Tk::LParen)) {
                advance();
                e->kind = AExpr::Call;
                e->name = name;
                if (!at(Tk::RParen)) {
                    e->args.push_back(expr());
                    while (at(Tk::Comma)) {
                        advance();
                        e->args.push_back(expr());
                    }
                }
                e->end = cur_end();
                expect(Tk::RParen);
                return e;
            }
            e->kind = AExpr::Ident;
            e->name = name;
            return e;
        }
        if (k && (k->kind == Tk::Notional || k->kind == Tk::Maturity)) {
            e->kind = AExpr::Ident;
            e->name = k->kind == Tk::Notional ? "notional" : "maturity";
            advance();
            return e;
        }
        parse_error("expected expression, got " + peek_debug(), s, e0);
    }

    // ---- statements (parser.rs:398-498) ----
    AStmt statement() {
        AStmt st;
        st.start = cur_start();
        const size_t s0 = cur_start(), e0 = cur_end();
        if (at(Tk::Let) || at(Tk::Set)) {
            st.kind = at(Tk::Let) ? AStmt::Let : AStmt::Set;
            advance();
            size_t a, b;
            st.name = expect_ident(a, b);
            expect(Tk::Eq);
            st.expr = expr();
        } else if (at(Tk::If)) {
            return if_statement();
        } else if (at(Tk::Pay) || at(Tk::Redeem)) {
            st.kind = at(Tk::Pay) ? AStmt::Pay : AStmt::Redeem;
            advance();
            st.expr = expr();
        } else if (at(Tk::Skip)) {
            st.kind = AStmt::Skip;
            advance();
        } else {
            parse_error("expected statement (let, if, pay, redeem, set, skip), got " + peek_debug(),
                        s0, e0);
        }
        st.end = cur_start();
        return st;
    }
    AStmt if_statement() {
        AStmt st;
        st.kind = AStmt::If;
        st.start = cur_start();
        expect(Tk::If);
        st.expr = expr();
        expect(Tk::Then);
        expect(Tk::Indent);
        while (!at_block_end()) st.then_body.push_back(statement());
        expect(Tk::Dedent);
        if (at(Tk::Else)) {
            advance();
            if (at(Tk::If)) {
                st.else_body.push_back(if_statement());
            } else {
                expect(Tk::Indent);
                while (!at_block_end()) st.else_body.push_back(statement());
                expect(Tk::Dedent);
            }
        }
        st.end = cur_start();
        return st;
    }

    // ---- product items (parser.rs:158-396) ----
    AItem item() {
        AItem it;
        it.start = cur_start();
        if (at(Tk::Notional) || at(Tk::Maturity)) {
            it.kind = at(Tk::Notional) ? AItem::Notional : AItem::Maturity;
            advance();
            expect(Tk::Colon);
            const Token &n = expect_number();
            it.value = n.num;
            it.end = n.end;
            return it;
        }
        if (at(Tk::Underlyings)) {
            it.kind = AItem::Underlyings;
            advance();
            expect(Tk::Indent);
            while (!at_block_end()) {
                AUnderlying u;
                u.name = expect_ident(u.start, u.end);
                u.index = (int64_t)it.und.size();
                u.type = UnderlyingType::Equity;
                if (at(Tk::Eq)) {
                    advance();
                    if (at(Tk::Asset) || at(Tk::Equity)) u.type = UnderlyingType::Equity;
                    else if (at(Tk::Fx)) u.type = UnderlyingType::Fx;
                    else if (at(Tk::Commodity)) u.type = UnderlyingType::Commodity;
                    else if (at(Tk::Rate)) u.type = UnderlyingType::Rate;
                    else
                        parse_error("expected asset type (asset, equity, fx, commodity, rate), got " +
                                        peek_debug(),
                                    cur_start(), cur_end());
                    advance();
                    expect(Tk::LParen);
                    const Token &n = expect_number();
                    // `idx as usize` saturates
                    const double v = n.num;
                    u.index = std::isnan(v) || v <= 0.0 ? 0
                              : v >= 9.2e18 ? INT64_MAX
                                            : (int64_t)v;
                    u.end = n.end;
                    expect(Tk::RParen);
                }
                it.und.push_back(std::move(u));
            }
            it.end = cur_end();
            expect(Tk::Dedent);
            return it;
        }
        if (at(Tk::State)) {
            it.kind = AItem::State;
            advance();
            expect(Tk::Indent);
            while (!at_block_end()) {
                AState d;
                size_t e;
                d.name = expect_ident(d.start, e);
                expect(Tk::Colon);
                if (at(Tk::Bool)) d.type_name = "bool";
                else if (at(Tk::Float)) d.type_name = "float";
                else if (at(Tk::Ident)) d.type_name = t[pos].text;
                else parse_error("expected type name (bool or float)", cur_start(), cur_end());
                advance();
                expect(Tk::Eq);
                d.init = expr();
                d.end = cur_start();
                it.state.push_back(std::move(d));
            }
            it.end = cur_end();
            expect(Tk::Dedent);
            return it;
        }
        if (at(Tk::Schedule)) {
            it.kind = AItem::Schedule;
            advance();
            const Token *k = peek();
            if (k && k->kind == Tk::Monthly) it.sched.period = 1.0 / 12.0;
            else if (k && k->kind == Tk::Quarterly) it.sched.period = 0.25;
            else if (k && k->kind == Tk::SemiAnnual) it.sched.period = 0.5;
            else if (k && k->kind == Tk::Annual) it.sched.period = 1.0;
            else if (k && k->kind == Tk::Number) it.sched.period = k->num;
            else parse_error("expected schedule frequency", cur_start(), cur_end());
            advance();
            expect(Tk::From);
            it.sched.start = expect_number().num;
            expect(Tk::To);
            it.sched.end = expect_number().num;
            expect(Tk::Indent);
            while (!at_block_end()) it.sched.body.push_back(statement());
            it.end = cur_end();
            expect(Tk::Dedent);
            return it;
        }
        if (pos < t.size())
            parse_error("unexpected token " + debug_kind(t[pos]) + " in product body", t[pos].start,
                        t[pos].end);
        parse_error("unexpected end of input in product body", eof_pos(), eof_pos());
    }

    AProduct product() {
        AProduct p;
        expect(Tk::Product);
        p.name = expect_string().text;
        expect(Tk::Indent);
        while (!at_block_end()) p.body.push_back(item());
        expect(Tk::Dedent);
        return p;
    }
};

// ScheduleFreq::generate_dates (ast.rs:64-80)
std::vector<double> generate_dates(double period, double start, double end) {
    if (period <= 0.0) return {end};
    std::vector<double> d;
    double t = start;
    while (t <= end + 1e-12) {
        d.push_back(t);
        t += period;
    }
    if (d.empty() || std::fabs(d.back() - end) > 1e-12) d.push_back(end);
    return d;
}

// ---------------------------------------------------------------------------
// Compiler (compiler.rs)
// ---------------------------------------------------------------------------

struct Scope {
    std::map<std::string, int> state, locals, underlyings;
    int next_local = 0;
};

std::unique_ptr<Expr> compile_expr(const AExpr &e, const Scope &sc) {
    auto out = std::make_unique<Expr>();
    switch (e.kind) {
    case AExpr::Num:
        out->kind = Expr::LitF64;
        out->num = e.num;
        return out;
    case AExpr::BoolLit:
        out->kind = Expr::LitBool;
        out->flag = e.b;
        return out;
    case AExpr::Ident: {   // resolve_ident (compiler.rs:277-296)
        auto l = sc.locals.find(e.name);
        if (l != sc.locals.end()) {
            out->kind = Expr::Local;
            out->slot = l->second;
            return out;
        }
        auto s = sc.state.find(e.name);
        if (s != sc.state.end()) {
            out->kind = Expr::State;
            out->slot = s->second;
            return out;
        }
        if (e.name == "notional") out->kind = Expr::Notional;
        else if (e.name == "observation_date") out->kind = Expr::Date;
        else if (e.name == "is_final") out->kind = Expr::IsFinal;
        else compile_error("undefined variable '" + e.name + "'", true, e.start, e.end);
        return out;
    }
    case AExpr::Bin:
        out->kind = Expr::Bin;
        out->bop = e.bop;
        out->args.push_back(compile_expr(*e.args[0], sc));
        out->args.push_back(compile_expr(*e.args[1], sc));
        return out;
    case AExpr::Un:
        out->kind = Expr::Un;
        out->uop = e.uop;
        out->args.push_back(compile_expr(*e.args[0], sc));
        return out;
    case AExpr::Call: {   // compile_fn_call (compiler.rs:298-330)
        static const std::map<std::string, Builtin> fns = {
            {"worst_of", Builtin::WorstOf}, {"best_of", Builtin::BestOf},
            {"performances", Builtin::Performances}, {"price", Builtin::Price},
            {"min", Builtin::Min}, {"max", Builtin::Max}, {"abs", Builtin::Abs},
            {"exp", Builtin::Exp}, {"log", Builtin::Log}};
        auto f = fns.find(e.name);
        if (f == fns.end()) compile_error("unknown function '" + e.name + "'", true, e.start, e.end);
        out->kind = Expr::Call;
        out->fn = f->second;
        for (const auto &a : e.args) out->args.push_back(compile_expr(*a, sc));
        return out;
    }
    }
    return out;
}

Stmt compile_stmt(const AStmt &s, Scope &sc) {
    Stmt o;
    switch (s.kind) {
    case AStmt::Let: {
        o.kind = Stmt::Let;
        o.expr = compile_expr(*s.expr, sc);
        o.slot = sc.next_local++;
        sc.locals[s.name] = o.slot;
        return o;
    }
    case AStmt::If:
        o.kind = Stmt::If;
        o.expr = compile_expr(*s.expr, sc);
        for (const auto &c : s.then_body) o.then_body.push_back(compile_stmt(c, sc));
        for (const auto &c : s.else_body) o.else_body.push_back(compile_stmt(c, sc));
        return o;
    case AStmt::Pay:
    case AStmt::Redeem:
        o.kind = s.kind == AStmt::Pay ? Stmt::Pay : Stmt::Redeem;
        o.expr = compile_expr(*s.expr, sc);
        return o;
    case AStmt::Set: {
        o.kind = Stmt::Set;
        auto it = sc.state.find(s.name);
        if (it == sc.state.end())
            compile_error("undefined state variable '" + s.name + "'", true, s.start, s.end);
        o.slot = it->second;
        o.expr = compile_expr(*s.expr, sc);
        return o;
    }
    case AStmt::Skip:
        o.kind = Stmt::Skip;
        return o;
    }
    return o;
}

std::string debug_ast_kind(const AExpr &e) {
    switch (e.kind) {
    case AExpr::Num: return "NumberLit(" + debug_f64(e.num) + ")";
    case AExpr::BoolLit: return std::string("BoolLit(") + (e.b ? "true" : "false") + ")";
    case AExpr::Ident: return "Ident(" + debug_str(e.name) + ")";
    case AExpr::Bin: return "BinOp { .. }";
    case AExpr::Un: return "UnaryOp { .. }";
    case AExpr::Call: return "FnCall { name: " + debug_str(e.name) + ", .. }";
    }
    return "?";
}

std::unique_ptr<Product> compile_product(const AProduct &ast) {
    auto p = std::make_unique<Product>();
    p->name = ast.name;
    Scope sc;
    bool has_notional = false, has_maturity = false;
    for (const AItem &it : ast.body) {
        switch (it.kind) {
        case AItem::Notional:
            if (has_notional) compile_error("duplicate notional declaration", true, it.start, it.end);
            has_notional = true;
            p->notional = it.value;
            break;
        case AItem::Maturity:
            if (has_maturity) compile_error("duplicate maturity declaration", true, it.start, it.end);
            has_maturity = true;
            p->maturity = it.value;
            break;
        case AItem::Underlyings:
            for (const AUnderlying &u : it.und) {
                if (sc.underlyings.count(u.name))
                    compile_error("duplicate underlying '" + u.name + "'", true, u.start, u.end);
                sc.underlyings[u.name] = (int)u.index;
                p->underlyings.push_back({u.name, u.index, u.type});
            }
            break;
        case AItem::State:
            for (const AState &d : it.state) {
                const int slot = (int)p->state_vars.size();
                if (sc.state.count(d.name))
                    compile_error("duplicate state variable '" + d.name + "'", true, d.start, d.end);
                // compile_initial_value (compiler.rs:152-166)
                StateDef def{d.name, 0.0, false};
                const AExpr &v = *d.init;
                if (v.kind == AExpr::BoolLit && d.type_name == "bool") {
                    def.initial = v.b ? 1.0 : 0.0;
                    def.is_bool = true;
                } else if (v.kind == AExpr::Num && d.type_name == "float") {
                    def.initial = v.num;
                } else if (v.kind == AExpr::Num && d.type_name == "bool") {
                    def.initial = v.num != 0.0 ? 1.0 : 0.0;
                    def.is_bool = true;
                } else if (v.kind == AExpr::BoolLit && d.type_name == "float") {
                    def.initial = v.b ? 1.0 : 0.0;
                } else {
                    compile_error("cannot use " + debug_ast_kind(v) + " as initial value for type '" +
                                      d.type_name + "'",
                                  true, v.start, v.end);
                }
                sc.state[d.name] = slot;
                p->state_vars.push_back(def);
            }
            break;
        case AItem::Schedule:
            break;
        }
    }
    if (!has_notional) compile_error("missing notional declaration", false, 0, 0);
    if (!has_maturity) compile_error("missing maturity declaration", false, 0, 0);
    for (const AItem &it : ast.body) {
        if (it.kind != AItem::Schedule) continue;
        sc.locals.clear();
        sc.next_local = 0;
        Schedule s;
        s.dates = generate_dates(it.sched.period, it.sched.start, it.sched.end);
        for (const AStmt &st : it.sched.body) s.body.push_back(compile_stmt(st, sc));
        p->schedules.push_back(std::move(s));
    }
    p->num_underlyings = std::max<int64_t>((int64_t)p->underlyings.size(), 1);
    return p;
}

// ---------------------------------------------------------------------------
// Device program builder
// ---------------------------------------------------------------------------

struct Builder {
    Program prog;
    int base = 0;        // first temporary register
    int depth = 0;       // current stack depth
    int max_depth = 0;
    int mask_depth = 0;

    int push() {
        const int r = base + depth++;
        max_depth = std::max(max_depth, depth);
        return r;
    }
    void pop(int k = 1) { depth -= k; }
    void emit(ofg::DslOp op, int dst, int a = 0, int b = 0, uint32_t aux = 0) {
        ofg::DslIns i;
        i.op = op;
        i.dst = (uint8_t)dst;
        i.a = (uint8_t)a;
        i.b = (uint8_t)b;
        i.aux = aux;
        prog.code.push_back(i);
    }
    uint32_t constant(double v) {   // add_constant (eval.rs:197-207): dedup by bits
        for (size_t i = 0; i < prog.consts.size(); ++i)
            if (std::memcmp(&prog.consts[i], &v, sizeof v) == 0) return (uint32_t)i;
        prog.consts.push_back(v);
        return (uint32_t)prog.consts.size() - 1;
    }

    // Leaves the value on a new stack register and returns it (compile_expr,
    // eval.rs:1599-1641). Locals and state are read through a copy onto the
    // stack, as the reference's PUSH_LOCAL / PUSH_STATE.
    int expr(const Expr &e, int n_locals) {
        switch (e.kind) {
        case Expr::LitF64: {
            const int r = push();
            emit(ofg::DOP_CONST, r, 0, 0, constant(e.num));
            return r;
        }
        case Expr::LitBool: {
            const int r = push();
            emit(ofg::DOP_CONST, r, 0, 0, constant(e.flag ? 1.0 : 0.0));
            return r;
        }
        case Expr::Local:
        case Expr::State: {
            const int src = e.kind == Expr::Local ? e.slot : n_locals + e.slot;
            const int r = push();
            emit(ofg::DOP_SET, r, src, 1);   // b = 1: unmasked copy
            return r;
        }
        case Expr::Notional: {
            const int r = push();
            emit(ofg::DOP_NOTIONAL, r);
            return r;
        }
        case Expr::Date: {
            const int r = push();
            emit(ofg::DOP_DATE, r);
            return r;
        }
        case Expr::IsFinal: {
            const int r = push();
            emit(ofg::DOP_ISFINAL, r);
            return r;
        }
        case Expr::Bin: {
            const int l = expr(*e.args[0], n_locals);
            const int r = expr(*e.args[1], n_locals);
            static const ofg::DslOp ops[] = {ofg::DOP_ADD, ofg::DOP_SUB, ofg::DOP_MUL, ofg::DOP_DIV,
                                             ofg::DOP_EQ,  ofg::DOP_NE,  ofg::DOP_LT,  ofg::DOP_LE,
                                             ofg::DOP_GT,  ofg::DOP_GE,  ofg::DOP_AND, ofg::DOP_OR};
            emit(ops[(int)e.bop], l, l, r);
            pop();
            return l;
        }
        case Expr::Un: {
            const int r = expr(*e.args[0], n_locals);
            emit(e.uop == UnOp::Neg ? ofg::DOP_NEG : ofg::DOP_NOT, r, r);
            return r;
        }
        case Expr::Call:
            return call(e, n_locals);
        }
        return 0;
    }

    // compile_builtin (eval.rs:1643-1745)
    int call(const Expr &e, int n_locals) {
        const auto &args = e.args;
        switch (e.fn) {
        case Builtin::Performances:
            eval_error("performances() cannot be used standalone; wrap in worst_of() or best_of()");
        case Builtin::WorstOf:
        case Builtin::BestOf: {
            const bool worst = e.fn == Builtin::WorstOf;
            if (args.size() == 1 && args[0]->kind == Expr::Call && args[0]->fn == Builtin::Performances) {
                const int r = push();
                emit(worst ? ofg::DOP_WOF_PERF : ofg::DOP_BOF_PERF, r);
                return r;
            }
            const int first = base + depth;
            for (const auto &a : args) expr(*a, n_locals);
            pop((int)args.size());
            const int r = push();
            emit(worst ? ofg::DOP_WORST : ofg::DOP_BEST, r, first, 0, (uint32_t)args.size());
            return r;
        }
        case Builtin::Price: {
            if (args.size() != 1) eval_error("price() requires an asset index");
            const int r = expr(*args[0], n_locals);
            emit(ofg::DOP_PRICE, r, r);
            return r;
        }
        case Builtin::Min:
        case Builtin::Max: {
            const bool mn = e.fn == Builtin::Min;
            if (args.size() != 2) eval_error(mn ? "min() requires 2 arguments" : "max() requires 2 arguments");
            const int l = expr(*args[0], n_locals);
            const int r = expr(*args[1], n_locals);
            emit(mn ? ofg::DOP_MIN : ofg::DOP_MAX, l, l, r);
            pop();
            return l;
        }
        case Builtin::Abs:
        case Builtin::Exp:
        case Builtin::Log: {
            const char *nm = e.fn == Builtin::Abs ? "abs" : e.fn == Builtin::Exp ? "exp" : "log";
            if (args.size() != 1) eval_error(std::string(nm) + "() requires 1 argument");
            const int r = expr(*args[0], n_locals);
            emit(e.fn == Builtin::Abs ? ofg::DOP_ABS : e.fn == Builtin::Exp ? ofg::DOP_EXP : ofg::DOP_LOG,
                 r, r);
            return r;
        }
        }
        return 0;
    }

    void stmt(const Stmt &s, int n_locals) {
        switch (s.kind) {
        case Stmt::Let:
        case Stmt::Set: {
            const int r = expr(*s.expr, n_locals);
            emit(ofg::DOP_SET, s.kind == Stmt::Let ? s.slot : n_locals + s.slot, r, 0);
            pop();
            break;
        }
        case Stmt::Pay:
        case Stmt::Redeem: {
            const int r = expr(*s.expr, n_locals);
            emit(s.kind == Stmt::Pay ? ofg::DOP_PAY : ofg::DOP_REDEEM, 0, r);
            pop();
            break;
        }
        case Stmt::Skip:
            emit(ofg::DOP_SKIP, 0);
            break;
        case Stmt::If: {
            const int c = expr(*s.expr, n_locals);
            const int k = mask_depth++;
            prog.n_masks = std::max(prog.n_masks, mask_depth);
            emit(ofg::DOP_IF, 0, c, 0, (uint32_t)k);
            pop();
            for (const Stmt &t : s.then_body) stmt(t, n_locals);
            if (!s.else_body.empty()) {
                emit(ofg::DOP_ELSE, 0, 0, 0, (uint32_t)k);
                for (const Stmt &t : s.else_body) stmt(t, n_locals);
            }
            emit(ofg::DOP_ENDIF, 0, 0, 0, (uint32_t)k);
            --mask_depth;
            break;
        }
        }
    }
};

int max_slot(const std::vector<Stmt> &ss) {
    int m = 0;
    for (const Stmt &s : ss) {
        if (s.kind == Stmt::Let) m = std::max(m, s.slot + 1);
        if (s.kind == Stmt::If) m = std::max({m, max_slot(s.then_body), max_slot(s.else_body)});
    }
    return m;
}

}  // namespace

int Product::max_local_slots() const {
    int m = 0;
    for (const Schedule &s : schedules) m = std::max(m, max_slot(s.body));
    return m;
}

std::unique_ptr<Product> parse_and_compile(const std::string &source) {
    Parser p;
    p.t = tokenize(source);
    const AProduct ast = p.product();
    return compile_product(ast);
}

std::vector<Program> build_programs(const Product &p, int &n_locals) {
    n_locals = p.max_local_slots();
    const int fixed = n_locals + (int)p.state_vars.size();
    std::vector<Program> out;
    for (const Schedule &s : p.schedules) {
        Builder b;
        b.base = fixed;
        for (const Stmt &st : s.body) b.stmt(st, n_locals);
        b.prog.n_regs = fixed + b.max_depth;
        out.push_back(std::move(b.prog));
    }
    return out;
}

}  // namespace ofg_dsl
