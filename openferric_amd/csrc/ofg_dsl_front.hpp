// Front end of the structured-product language (src/dsl/lexer.rs, parser.rs,
// compiler.rs, ir.rs): source text -> tokens -> syntax tree -> compiled
// product (slots resolved, schedule dates generated), and the compiled
// product -> device program of ofg_dsl.hip. Host only.
#pragma once

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "ofg_dsl.hpp"

namespace ofg_dsl {

// DslError (src/dsl/error.rs:17-46); what() is the reference's Display text.
struct DslError : std::runtime_error {
    enum Kind { Lex, Parse, Compile, Eval } kind;
    DslError(Kind k, const std::string &m) : std::runtime_error(m), kind(k) {}
};

enum class BinOp { Add, Sub, Mul, Div, Eq, Ne, Lt, Le, Gt, Ge, And, Or };
enum class UnOp { Neg, Not };
enum class Builtin { WorstOf, BestOf, Performances, Price, Min, Max, Abs, Exp, Log };

// ir::Expr (src/dsl/ir.rs:93-116)
struct Expr {
    enum Kind { LitF64, LitBool, Local, State, Notional, Date, IsFinal, Bin, Un, Call } kind;
    double num = 0.0;
    bool flag = false;
    int slot = 0;
    BinOp bop = BinOp::Add;
    UnOp uop = UnOp::Neg;
    Builtin fn = Builtin::Abs;
    std::vector<std::unique_ptr<Expr>> args;   // Bin: lhs, rhs; Un: operand; Call: arguments
};

// ir::Statement (src/dsl/ir.rs:118-133)
struct Stmt {
    enum Kind { Let, If, Pay, Redeem, Set, Skip } kind;
    int slot = 0;
    std::unique_ptr<Expr> expr;   // Let/Pay/Redeem/Set value, If condition
    std::vector<Stmt> then_body, else_body;
};

struct Schedule {
    std::vector<double> dates;
    std::vector<Stmt> body;
};

enum class UnderlyingType { Equity, Fx, Commodity, Rate };

struct UnderlyingDef {
    std::string name;
    int64_t asset_index;
    UnderlyingType type;
};

struct StateDef {
    std::string name;
    double initial;   // Value::as_f64
    bool is_bool;
};

// ir::CompiledProduct (src/dsl/ir.rs:141-151)
struct Product {
    std::string name;
    double notional = 0.0, maturity = 0.0;
    int64_t num_underlyings = 1;
    std::vector<UnderlyingDef> underlyings;
    std::vector<StateDef> state_vars;
    std::vector<Schedule> schedules;
    int max_local_slots() const;
};

// parse_and_compile (src/dsl/mod.rs:24-28); throws DslError.
std::unique_ptr<Product> parse_and_compile(const std::string &source);

// One schedule's device program (the reference's bytecode compile_statement
// / compile_expr, eval.rs:1542-1772, as predicated register code).
struct Program {
    std::vector<ofg::DslIns> code;
    std::vector<double> consts;
    int n_regs = 0;      // locals + state + temporaries
    int n_masks = 0;     // if-nesting depth
};

// Compiles every schedule; registers [0, locals) hold the locals,
// [locals, locals + state) the state variables. Throws DslError (Eval) for
// the reference's build_execution_plan errors.
std::vector<Program> build_programs(const Product &p, int &n_locals);

}  // namespace ofg_dsl
