// tape.cc — prefix program -> Sethi-Ullman accumulator tape (see tape.h).
#include "tape.h"

#include <algorithm>
#include <cstring>
#include <thread>

namespace ppi {

void SymbolTable::build(const ppi_symbol* t, int32_t n) {
    entries.assign(t, t + n);
    direct.assign(8192, 0);
    other.clear();
    for (int32_t k = 0; k < n; ++k) {
        int32_t s = t[k].symbol - 10000;
        if (s >= 0 && s < 8192) {
            if (!direct[s]) direct[s] = k + 1;
        } else {
            other.emplace_back(t[k].symbol, k);
        }
    }
    std::stable_sort(other.begin(), other.end(),
                     [](const std::pair<int32_t, int32_t>& a, const std::pair<int32_t, int32_t>& b) {
                         return a.first < b.first;
                     });
}

const ppi_symbol* SymbolTable::find(int32_t symbol) const {
    int32_t s = symbol - 10000;
    if (s >= 0 && s < 8192) return direct[s] ? &entries[direct[s] - 1] : nullptr;
    auto it = std::lower_bound(other.begin(), other.end(), std::make_pair(symbol, INT32_MIN));
    if (it != other.end() && it->first == symbol) return &entries[it->second];
    return nullptr;
}

namespace {

struct Node {
    int32_t op;       // ppi_op of an operator, or -1 for a leaf
    int32_t ch[3];
    uint32_t arg;     // leaf: float bits or column
    int8_t leaf_attr; // leaf: 1 = attribute column, 0 = constant
    int8_t chain_ops; // leaf: operators of the chain in `chain`
    int32_t pos;      // phenotype position (for T_CONST bookkeeping)
    bool is_tconst;
    int16_t regs;
    uint64_t chain;   // leaf: chain key of an input or derived column (DerivedColumns), 0 = none
    int32_t fold;     // leaf: 1 + index of the operator node of the constant subtree it stands for, 0 = none
};

// Chain key of unary operator u applied to the chain `key` of `ops` operators.
inline uint64_t chain_extend(uint64_t key, int ops, uint32_t u) {
    return key | ((uint64_t)(u + 1) << (16 + 5 * ops));
}

uint32_t fbits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }

struct Compiler {
    const SymbolTable& st;
    int32_t ncol;
    bool want_const_pos;
    const DerivedColumns* derived = nullptr;
    int64_t derived_uses = 0;
    bool fold_consts = false;  // constant subtrees go to Tape::folds
    bool fold_too_deep = false;
    int64_t folded_nodes = 0;
    std::vector<std::pair<int32_t, uint32_t>> fold_pending;  // (subtree root, placeholder instruction)
    Tape fold_scratch;
    std::vector<Node> nodes;
    std::vector<int32_t> vstack;
    struct Frame { int32_t node, base, phase, o[3]; };
    std::vector<Frame> frames;

    Compiler(const SymbolTable& s, int32_t nc, bool w) : st(s), ncol(nc), want_const_pos(w) {}

    // Returns slots used, -1 if invalid (underflow / empty stack).
    int compile(const int32_t* ph, const float* eph, int32_t size, Tape* out) {
        nodes.clear();
        vstack.clear();
        for (int32_t i = size - 1; i >= 0; --i) {
            const ppi_symbol* s = st.find(ph[i]);
            int32_t op = s ? s->op : PPI_OP_UNKNOWN;
            int ar = arity(op);
            Node n{};
            n.pos = i;
            n.regs = 0;
            n.chain = 0;
            n.chain_ops = 0;
            if (ar == 0) {
                n.op = -1;
                n.leaf_attr = 0;
                n.is_tconst = false;
                if (op == PPI_OP_CONST) {
                    n.arg = fbits(eph[i]);
                    n.is_tconst = true;
                } else if (op == PPI_OP_ATTRIBUTE) {
                    float c = eph[i];
                    if (c > -1.0f && c < (float)ncol) {
                        n.leaf_attr = 1;
                        n.arg = (uint32_t)(int32_t)c;
                        if (ncol < 65535) n.chain = (uint64_t)n.arg + 1;
                    } else {
                        n.arg = fbits(__builtin_nanf(""));
                    }
                } else if (op == PPI_OP_VALUE) {
                    n.arg = fbits(s->value);
                } else {
                    n.arg = fbits(__builtin_nanf(""));
                }
            } else {
                if ((int)vstack.size() < ar) return -1;
                if (ar == 1 && derived) {
                    // Unary operator over a column (an input column or a derived one):
                    // becomes a leaf reading the derived column of the extended chain,
                    // if the plan gave it one.
                    Node& leaf = nodes[(size_t)vstack.back()];
                    if (leaf.op < 0 && leaf.leaf_attr && leaf.chain && leaf.chain_ops < kChainMaxOps) {
                        const uint64_t key = chain_extend(leaf.chain, leaf.chain_ops, unop_of(op));
                        const int32_t dcol = derived->find(key);
                        if (dcol >= 0) {
                            leaf.arg = (uint32_t)dcol;
                            leaf.chain = key;
                            leaf.chain_ops++;
                            leaf.pos = i;
                            ++derived_uses;
                            continue;
                        }
                    }
                }
                n.op = op;
                for (int k = 0; k < ar; ++k) { n.ch[k] = vstack.back(); vstack.pop_back(); }
                bool constant = fold_consts;
                for (int k = 0; k < ar && constant; ++k) {
                    const Node& c = nodes[(size_t)n.ch[k]];
                    constant = c.op < 0 && !c.leaf_attr;
                }
                if (constant) {
                    // Every operand is a constant: the operator node keeps its
                    // children (a folded child contributes its whole subtree) and a
                    // constant leaf standing for its value takes its place.
                    for (int k = 0; k < ar; ++k)
                        if (nodes[(size_t)n.ch[k]].fold) n.ch[k] = nodes[(size_t)n.ch[k]].fold - 1;
                }
                // Sethi-Ullman register need.
                if (ar == 1) {
                    n.regs = nodes[n.ch[0]].regs;
                } else if (ar == 2) {
                    const Node& a = nodes[n.ch[0]];
                    const Node& b = nodes[n.ch[1]];
                    bool la = a.op < 0, lb = b.op < 0;
                    if (la && lb) n.regs = 0;
                    else if (la) n.regs = b.regs;
                    else if (lb) n.regs = a.regs;
                    else n.regs = (int16_t)(a.regs == b.regs ? a.regs + 1 : std::max(a.regs, b.regs));
                } else {
                    int16_t r[3] = {nodes[n.ch[0]].regs, nodes[n.ch[1]].regs, nodes[n.ch[2]].regs};
                    std::sort(r, r + 3, [](int16_t x, int16_t y) { return x > y; });
                    n.regs = (int16_t)std::max<int>(r[0], std::max<int>(1 + r[1], 2 + r[2]));
                }
                if (constant) {
                    nodes.push_back(n);
                    Node leaf{};
                    leaf.op = -1;
                    leaf.pos = i;
                    leaf.fold = (int32_t)nodes.size();
                    nodes.push_back(leaf);
                    vstack.push_back((int32_t)nodes.size() - 1);
                    continue;
                }
            }
            nodes.push_back(n);
            vstack.push_back((int32_t)nodes.size() - 1);
        }
        if (vstack.empty()) return -1;
        root_node = vstack.back();
        fold_pending.clear();
        const int slots = emit(root_node, out);
        if (!fold_pending.empty()) emit_folds(out);
        return slots;
    }

    // The constant subtrees the program's tape refers to, each compiled into a
    // program of its own in out->folds.
    void emit_folds(Tape* out) {
        FoldSet& F = out->folds;
        const auto pending = fold_pending;
        for (const auto& fp : pending) {
            fold_scratch.code.clear();
            const int slots = emit(fp.first, &fold_scratch);
            if (slots > kMaxSlots) fold_too_deep = true;
            ProgMeta m{};
            m.code_off = (uint32_t)F.code.size();
            m.ncode = (uint32_t)fold_scratch.code.size();
            m.flags = (uint32_t)slots << 8;
            for (const Ins& in : fold_scratch.code) {
                const uint32_t h = in.code & 0xffu;
                folded_nodes += (h >= H_UN || h == H_TERN) ? 1 : 0;
            }
            F.code.insert(F.code.end(), fold_scratch.code.begin(), fold_scratch.code.end());
            F.code.push_back(Ins{H_END, 0});
            F.meta.push_back(m);
            F.at.push_back(fp.second);
        }
        fold_pending.clear();
    }

    int pending_save = -1;
    int max_slot = 0;
    int32_t root_node = -1;

    // Derivative bookkeeping (compile_adjoint): the tape index of each operator
    // node's instruction, and the operand side (0 = accumulator, 1 = folded or
    // parked operand) on which each node enters its parent's instruction.
    bool track = false;
    std::vector<int32_t> ins_at;
    std::vector<uint8_t> side;

    void put(Tape* out, uint32_t code, uint32_t arg, uint32_t cpos) {
        out->code.push_back(Ins{code, arg});
        if (want_const_pos) out->const_pos.push_back(cpos);
    }
    void mark(int32_t node, Tape* out) {
        if (track) ins_at[node] = (int32_t)out->code.size() - 1;
    }
    void mark_side(int32_t node, uint8_t s) {
        if (track) side[node] = s;
    }

    // A leaf's argument; a folded subtree's leaf gets a placeholder that
    // apply_folds() replaces with the subtree's value.
    uint32_t leaf_arg(const Node& n, Tape* out) {
        if (!n.fold) return n.arg;
        fold_pending.emplace_back(n.fold - 1, (uint32_t)out->code.size());
        return 0;
    }

    void emit_load(const Node& n, Tape* out) {
        uint32_t cpos = n.is_tconst ? (uint32_t)n.pos + 1 : 0;
        if (pending_save >= 0) {
            uint32_t h = n.leaf_attr ? H_PUSH_A : H_PUSH_C;
            put(out, h | ((uint32_t)pending_save << 8), leaf_arg(n, out), cpos);
            max_slot = std::max(max_slot, pending_save + 1);
            pending_save = -1;
        } else {
            put(out, n.leaf_attr ? H_LOAD_A : H_LOAD_C, leaf_arg(n, out), cpos);
        }
    }

    void emit_fold(uint32_t bop, const Node& leaf, Tape* out) {
        uint32_t cpos = leaf.is_tconst ? (uint32_t)leaf.pos + 1 : 0;
        put(out, (leaf.leaf_attr ? H_BIN_A : H_BIN_C) + bop, leaf_arg(leaf, out), cpos);
    }

    int emit(int32_t root, Tape* out) {
        pending_save = -1;
        max_slot = 0;
        if (track) {
            ins_at.assign(nodes.size(), -1);
            side.assign(nodes.size(), 0);
        }
        frames.clear();
        frames.push_back(Frame{root, 0, 0, {0, 0, 0}});
        while (!frames.empty()) {
            Frame& f = frames.back();
            const Node& n = nodes[f.node];
            if (n.op < 0) {
                emit_load(n, out);
                frames.pop_back();
                continue;
            }
            int ar = arity(n.op);
            if (ar == 1) {
                if (f.phase == 0) {
                    f.phase = 1;
                    int32_t c = n.ch[0], b = f.base;
                    frames.push_back(Frame{c, b, 0, {0, 0, 0}});
                } else {
                    put(out, H_UN + unop_of(n.op), 0, 0);
                    mark(f.node, out);
                    mark_side(n.ch[0], 0);
                    frames.pop_back();
                }
                continue;
            }
            if (ar == 2) {
                const Node& L = nodes[n.ch[0]];
                const Node& R = nodes[n.ch[1]];
                bool lleaf = L.op < 0, rleaf = R.op < 0;
                if (lleaf && rleaf) {
                    emit_load(L, out);
                    emit_fold(binop_of(n.op, false), R, out);
                    mark(f.node, out);
                    mark_side(n.ch[0], 0);
                    mark_side(n.ch[1], 1);
                    frames.pop_back();
                } else if (lleaf || rleaf) {
                    if (f.phase == 0) {
                        f.phase = 1;
                        int32_t c = lleaf ? n.ch[1] : n.ch[0], b = f.base;
                        frames.push_back(Frame{c, b, 0, {0, 0, 0}});
                    } else {
                        // acc holds the computed child; fold the leaf on its side.
                        emit_fold(binop_of(n.op, /*rev=*/lleaf), lleaf ? L : R, out);
                        mark(f.node, out);
                        mark_side(n.ch[0], lleaf ? 1 : 0);
                        mark_side(n.ch[1], lleaf ? 0 : 1);
                        frames.pop_back();
                    }
                } else {
                    bool left_first = L.regs >= R.regs;
                    if (f.phase == 0) {
                        f.phase = 1;
                        int32_t c = left_first ? n.ch[0] : n.ch[1], b = f.base;
                        frames.push_back(Frame{c, b, 0, {0, 0, 0}});
                    } else if (f.phase == 1) {
                        f.phase = 2;
                        pending_save = f.base;
                        int32_t c = left_first ? n.ch[1] : n.ch[0], b = f.base + 1;
                        frames.push_back(Frame{c, b, 0, {0, 0, 0}});
                    } else {
                        // r[base] holds the first-evaluated child, acc the second.
                        put(out, H_BIN_S + binop_of(n.op, /*rev=*/left_first) + ((uint32_t)f.base << 8), 0, 0);
                        mark(f.node, out);
                        mark_side(n.ch[0], left_first ? 1 : 0);
                        mark_side(n.ch[1], left_first ? 0 : 1);
                        frames.pop_back();
                    }
                }
                continue;
            }
            // ternary: IF_THEN_ELSE(cond=ch0, then=ch1, else=ch2)
            if (f.phase == 0) {
                int32_t ord[3] = {0, 1, 2};
                std::stable_sort(ord, ord + 3, [&](int32_t x, int32_t y) {
                    return nodes[n.ch[x]].regs > nodes[n.ch[y]].regs;
                });
                for (int k = 0; k < 3; ++k) f.o[k] = ord[k];
                f.phase = 1;
                int32_t c = n.ch[ord[0]], b = f.base;
                frames.push_back(Frame{c, b, 0, {0, 0, 0}});
            } else if (f.phase == 1) {
                f.phase = 2;
                pending_save = f.base;
                int32_t c = n.ch[f.o[1]], b = f.base + 1;
                frames.push_back(Frame{c, b, 0, {0, 0, 0}});
            } else if (f.phase == 2) {
                f.phase = 3;
                pending_save = f.base + 1;
                int32_t c = n.ch[f.o[2]], b = f.base + 2;
                frames.push_back(Frame{c, b, 0, {0, 0, 0}});
            } else {
                // where[role]: 1 = r[slot] (first), 2 = r[slot2] (second), 0 = acc (third)
                uint32_t where[3];
                where[f.o[0]] = 1;
                where[f.o[1]] = 2;
                where[f.o[2]] = 0;
                uint32_t code = H_TERN | ((uint32_t)f.base << 8) | ((uint32_t)(f.base + 1) << 13) |
                                (where[0] << 18) | (where[1] << 20) | (where[2] << 22);
                put(out, code, 0, 0);
                mark(f.node, out);
                frames.pop_back();
            }
        }
        return max_slot;
    }
};

void put_end(Tape* out, bool want_const_pos, uint32_t flags) {
    out->code.push_back(Ins{H_END, flags});
    if (want_const_pos) out->const_pos.push_back(0);
}

}  // namespace

namespace {

// Chains of unary operators over input columns in programs [lo, hi) (every `step`-th program),
// counted into `counts`.
void count_chains(const SymbolTable& st, const int32_t* phenotype, const float* ephemeral, const int32_t* size,
                  int32_t stride, int32_t lo, int32_t hi, int32_t ncol, int32_t step, FlatMap64<uint32_t>* counts) {
    struct Item { uint64_t key; int ops; };
    std::vector<Item> stk;
    for (int32_t p = lo; p < hi; ++p) {
        if (p % step) continue;
        const int32_t* ph = phenotype + (size_t)p * stride;
        const float* eph = ephemeral + (size_t)p * stride;
        stk.clear();
        for (int32_t i = size[p] - 1; i >= 0; --i) {
            const ppi_symbol* s = st.find(ph[i]);
            const int32_t op = s ? s->op : PPI_OP_UNKNOWN;
            const int ar = arity(op);
            if (ar == 0) {
                const float c = eph[i];
                const bool col = op == PPI_OP_ATTRIBUTE && c > -1.0f && c < (float)ncol && ncol < 65535;
                stk.push_back(Item{col ? (uint64_t)(int32_t)c + 1 : 0, 0});
            } else if ((int)stk.size() < ar) {
                break;  // underflow: the compiler rejects the program
            } else if (ar == 1) {
                Item it = stk.back();
                stk.pop_back();
                if (it.key && it.ops < kChainMaxOps) {
                    it.key = chain_extend(it.key, it.ops, unop_of(op));
                    it.ops++;
                    ++counts->at(it.key);
                } else {
                    it.key = 0;
                }
                stk.push_back(it);
            } else {
                stk.resize(stk.size() - (size_t)ar);
                stk.push_back(Item{0, 0});
            }
        }
    }
}

int chain_ops_of(uint64_t key) {
    int ops = 0;
    while (ops < kChainMaxOps && ((key >> (16 + 5 * ops)) & 31u)) ++ops;
    return ops;
}

bool compile_range(const SymbolTable& st, const int32_t* phenotype, const float* ephemeral, const int32_t* size,
                   int32_t stride, int32_t n, int32_t ncol, bool want_const_pos, Tape* out, std::string* err,
                   const DerivedColumns* derived, bool fold);

}  // namespace

void apply_folds(Tape* t, const float* value) {
    for (size_t j = 0; j < t->folds.at.size(); ++j) t->code[t->folds.at[j]].arg = fbits(value[j]);
}

void plan_derived(const SymbolTable& st, const int32_t* phenotype, const float* ephemeral, const int32_t* size,
                  int32_t stride, int32_t n, int32_t ncol, int threads, DerivedColumns* derived) {
    if (!derived || derived->cap <= 0) return;
    const bool full = (int32_t)derived->cols.size() >= derived->cap;
    if (full && !derived->refill) return;
    threads = std::max(1, std::min(threads, n / 128));
    // Chain counts of every `step`-th program, on `threads` workers, merged in worker order.
    auto count_all = [&](int32_t step) {
        std::vector<FlatMap64<uint32_t>> counts((size_t)threads);
        if (threads == 1) {
            count_chains(st, phenotype, ephemeral, size, stride, 0, n, ncol, step, &counts[0]);
        } else {
            std::vector<std::thread> pool;
            const int32_t chunk = (n + threads - 1) / threads;
            for (int t = 0; t < threads; ++t) {
                const int32_t lo = t * chunk, hi = std::min(n, lo + chunk);
                if (lo < hi)
                    pool.emplace_back([&, t, lo, hi] {
                        count_chains(st, phenotype, ephemeral, size, stride, lo, hi, ncol, step, &counts[(size_t)t]);
                    });
            }
            for (auto& th : pool) th.join();
            for (int t = 1; t < threads; ++t)
                counts[(size_t)t].for_each([&](uint64_t k, uint32_t v) { counts[0].at(k) += v; });
        }
        return std::move(counts[0]);
    };
    if (full) {
        // Is the table still used?  Judged on every 8th program of a large population (the
        // share of chain occurrences the table covers; a full count follows only a refill).
        const FlatMap64<uint32_t> sample = count_all(n >= 1024 ? 8 : 1);
        uint64_t all = 0, covered = 0;
        sample.for_each([&](uint64_t k, uint32_t v) {
            all += v;
            if (derived->index.contains(k)) covered += v;
        });
        if (all == 0 || (double)covered >= kDerivedKeepShare * (double)all) return;
        derived->reset(derived->base, derived->cap);
    }
    const FlatMap64<uint32_t> counts = count_all(1);
    struct Cand { uint64_t key; uint32_t count; int ops; };
    std::vector<Cand> cand;
    cand.reserve(counts.size());
    counts.for_each([&](uint64_t k, uint32_t v) {
        if (!derived->index.contains(k)) cand.push_back(Cand{k, v, chain_ops_of(k)});
    });
    std::sort(cand.begin(), cand.end(), [](const Cand& a, const Cand& b) {
        if (a.count != b.count) return a.count > b.count;
        if (a.ops != b.ops) return a.ops < b.ops;
        return a.key < b.key;
    });
    for (const Cand& c : cand) {
        if ((int32_t)derived->cols.size() >= derived->cap) break;
        const int last = c.ops - 1;
        const uint32_t u = (uint32_t)((c.key >> (16 + 5 * last)) & 31u) - 1u;
        int32_t src;
        if (last == 0) {
            src = (int32_t)(c.key & 0xffffu) - 1;
        } else {
            const uint64_t prefix = c.key & ~((uint64_t)31 << (16 + 5 * last));
            src = derived->find(prefix);
            if (src < 0) continue;  // the prefix got no column
        }
        derived->index.at(c.key) = derived->base + (int32_t)derived->cols.size();
        derived->cols.push_back(DerivedColumns::Col{u, src});
    }
}

bool compile_population_planned(const SymbolTable& st, const int32_t* phenotype, const float* ephemeral,
                                const int32_t* size, int32_t stride, int32_t n, int32_t ncol, Tape* out,
                                std::string* err, const DerivedColumns* derived, bool fold) {
    return compile_range(st, phenotype, ephemeral, size, stride, n, ncol, false, out, err, derived, fold);
}

bool compile_population(const SymbolTable& st, const int32_t* phenotype, const float* ephemeral,
                        const int32_t* size, int32_t stride, int32_t n, int32_t ncol, bool want_const_pos,
                        Tape* out, std::string* err, DerivedColumns* derived, bool fold) {
    if (derived) plan_derived(st, phenotype, ephemeral, size, stride, n, ncol, 1, derived);
    return compile_range(st, phenotype, ephemeral, size, stride, n, ncol, want_const_pos, out, err, derived, fold);
}

namespace {
bool compile_range(const SymbolTable& st, const int32_t* phenotype, const float* ephemeral, const int32_t* size,
                   int32_t stride, int32_t n, int32_t ncol, bool want_const_pos, Tape* out, std::string* err,
                   const DerivedColumns* derived, bool fold) {
    Compiler c(st, ncol, want_const_pos);
    c.derived = derived;
    c.fold_consts = fold && !want_const_pos;
    out->meta.reserve(out->meta.size() + (size_t)n);
    for (int32_t i = 0; i < n; ++i) {
        ProgMeta m{};
        m.code_off = (uint32_t)out->code.size();
        m.size = size[i];
        int32_t sz = size[i];
        if (sz < 0 || sz > stride) {
            if (err) *err = "program " + std::to_string(i) + ": size " + std::to_string(sz) +
                            " outside [0, " + std::to_string(stride) + "]";
            return false;
        }
        if (sz == 0) {
            m.flags = kEmpty;
            m.ncode = 0;
            put_end(out, want_const_pos, kEmpty);
            out->stats.n_empty++;
            out->meta.push_back(m);
            continue;
        }
        int slots = c.compile(phenotype + (size_t)i * stride, ephemeral + (size_t)i * stride, sz, out);
        if (slots < 0) {
            // Underflowing program (undefined in the reference): evaluates to NaN.
            out->code.resize(m.code_off);
            if (want_const_pos) out->const_pos.resize(m.code_off);
            out->code.push_back(Ins{H_LOAD_C, 0x7fc00000u});
            if (want_const_pos) out->const_pos.push_back(0);
            m.flags = kInvalid;
            slots = 0;
            out->stats.n_invalid++;
        }
        if (slots > kMaxSlots || c.fold_too_deep) {
            if (err) *err = "program " + std::to_string(i) + " needs more than " + std::to_string(kMaxSlots) +
                            " register slots";
            return false;
        }
        m.ncode = (uint32_t)out->code.size() - m.code_off;
        put_end(out, want_const_pos, 0);
        m.flags |= (uint32_t)slots << 8;
        out->stats.total_size += sz;
        out->stats.total_code += m.ncode;
        out->stats.max_slots = std::max(out->stats.max_slots, slots);
        out->stats.max_code = std::max<int32_t>(out->stats.max_code, (int32_t)m.ncode);
        out->meta.push_back(m);
    }
    for (size_t k = 0; k < out->code.size(); ++k) {
        uint32_t h = out->code[k].code & 0xffu;
        if ((h == H_LOAD_A || h == H_PUSH_A || (h >= H_BIN_A && h < H_BIN_S)) && (int32_t)out->code[k].arg < ncol)
            out->stats.max_col = std::max<int32_t>(out->stats.max_col, (int32_t)out->code[k].arg);
    }
    out->stats.derived_uses += c.derived_uses;
    out->stats.folded_nodes += c.folded_nodes;
    return true;
}
}  // namespace

bool compile_population_mt(const SymbolTable& st, const int32_t* phenotype, const float* ephemeral,
                           const int32_t* size, int32_t stride, int32_t n, int32_t ncol, bool want_const_pos,
                           int threads, Tape* out, std::string* err, DerivedColumns* derived, bool fold) {
    out->meta.clear();
    out->code.clear();
    out->const_pos.clear();
    out->stats = CompileStats{};
    out->folds = FoldSet{};
    if (derived) plan_derived(st, phenotype, ephemeral, size, stride, n, ncol, threads, derived);
    // At least 128 programs per worker: a worker's start-up costs about as much as compiling
    // a few dozen programs.
    threads = std::min(threads, n / 128);
    if (threads <= 1)
        return compile_range(st, phenotype, ephemeral, size, stride, n, ncol, want_const_pos, out, err, derived,
                             fold);
    std::vector<Tape> parts((size_t)threads);
    std::vector<std::string> errs((size_t)threads);
    std::vector<char> ok((size_t)threads, 1);
    std::vector<std::thread> pool;
    int32_t chunk = (n + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
        int32_t lo = t * chunk, hi = std::min(n, lo + chunk);
        if (lo >= hi) continue;
        pool.emplace_back([&, t, lo, hi] {
            ok[t] = compile_range(st, phenotype + (size_t)lo * stride, ephemeral + (size_t)lo * stride, size + lo,
                                  stride, hi - lo, ncol, want_const_pos, &parts[t], &errs[t], derived, fold);
        });
    }
    for (auto& th : pool) th.join();
    for (int t = 0; t < threads; ++t)
        if (!ok[t]) {
            if (err) *err = errs[t] + " (in chunk starting at " + std::to_string(t * chunk) + ")";
            return false;
        }
    size_t ncode = 0;
    for (auto& p : parts) ncode += p.code.size();
    out->code.reserve(ncode);
    out->meta.reserve((size_t)n);
    for (auto& p : parts) {
        uint32_t base = (uint32_t)out->code.size();
        const uint32_t fbase = (uint32_t)out->folds.code.size();
        for (auto m : p.folds.meta) { m.code_off += fbase; out->folds.meta.push_back(m); }
        out->folds.code.insert(out->folds.code.end(), p.folds.code.begin(), p.folds.code.end());
        for (uint32_t a : p.folds.at) out->folds.at.push_back(a + base);
        for (auto m : p.meta) { m.code_off += base; out->meta.push_back(m); }
        out->code.insert(out->code.end(), p.code.begin(), p.code.end());
        if (want_const_pos) out->const_pos.insert(out->const_pos.end(), p.const_pos.begin(), p.const_pos.end());
        out->stats.total_size += p.stats.total_size;
        out->stats.total_code += p.stats.total_code;
        out->stats.max_slots = std::max(out->stats.max_slots, p.stats.max_slots);
        out->stats.max_code = std::max(out->stats.max_code, p.stats.max_code);
        out->stats.n_invalid += p.stats.n_invalid;
        out->stats.n_empty += p.stats.n_empty;
        out->stats.max_col = std::max(out->stats.max_col, p.stats.max_col);
        out->stats.derived_uses += p.stats.derived_uses;
        out->stats.folded_nodes += p.stats.folded_nodes;
    }
    return true;
}

// ---------------------------------------------------------------------------
// adjoint tape
// ---------------------------------------------------------------------------
namespace {

// Partial of an operator instruction with respect to one operand: 0 = zero
// everywhere it exists (the path is cut), 1 = the constant *v, 2 = stored by the
// forward pass.  side: 0 = accumulator operand, 1 = the other operand; for
// IF_THEN_ELSE the caller asks only about the then/else operands.
int partial_kind(uint32_t code, int side, float* v) {
    const uint32_t h = code & 0xffu;
    if (h == H_TERN) return 2;
    if (h >= H_UN && h < H_BIN_C) {
        switch (h - H_UN) {
        case U_NEG: *v = -1.0f; return 1;
        case U_NOT: case U_ROUND: case U_CEIL: case U_FLOOR: case U_STEP: case U_SIGN: return 0;
        default: return 2;
        }
    }
    uint32_t op = h >= H_BIN_S ? h - H_BIN_S : (h >= H_BIN_A ? h - H_BIN_A : h - H_BIN_C);
    switch (op) {
    case B_AND: case B_OR: case B_XOR: case B_GT: case B_GE: case B_LT: case B_LE: case B_EQ: case B_NE:
        return 0;
    case B_ADD: *v = 1.0f; return 1;
    case B_SUB: *v = side == 0 ? 1.0f : -1.0f; return 1;
    case B_RSUB: *v = side == 0 ? -1.0f : 1.0f; return 1;
    case B_MEAN: *v = 0.5f; return 1;
    case B_MOD: if (side == 0) { *v = 1.0f; return 1; } return 2;
    case B_RMOD: if (side == 1) { *v = 1.0f; return 1; } return 2;
    default: return 2;
    }
}

// Builds the jobs of one compiled program (Compiler state: nodes, ins_at, side, root_node).
struct AdjBuilder {
    const Compiler& c;
    const Tape& t;
    int32_t B;
    AdjTape* out;
    std::vector<uint8_t> live, act;
    std::vector<int32_t> need, kidx;
    struct Edge { int32_t parent, child; int kind; float v; };

    AdjBuilder(const Compiler& cc, const Tape& tt, int32_t b, AdjTape* o) : c(cc), t(tt), B(b), out(o) {}

    // Kind of the edge parent -> child (child index j in the tree).
    int edge_kind(int32_t parent, int j, float* v) const {
        const Node& n = c.nodes[parent];
        const uint32_t code = t.code[(size_t)c.ins_at[parent]].code;
        if (arity(n.op) == 3) return j == 0 ? 0 : 2;
        return partial_kind(code, c.side[n.ch[j]], v);
    }

    // Live T_CONST leaves of the root's tree (uncut path), in phenotype order.
    void live_consts(std::vector<int32_t>* consts) {
        const size_t nn = c.nodes.size();
        live.assign(nn, 0);
        std::vector<int32_t> st{c.root_node};
        live[(size_t)c.root_node] = 1;
        while (!st.empty()) {
            int32_t v = st.back();
            st.pop_back();
            const Node& n = c.nodes[(size_t)v];
            if (n.op < 0) {
                if (n.is_tconst) consts->push_back(v);
                continue;
            }
            for (int j = 0; j < arity(n.op); ++j) {
                float dummy;
                if (edge_kind(v, j, &dummy) == 0) continue;
                live[(size_t)n.ch[j]] = 1;
                st.push_back(n.ch[j]);
            }
        }
        std::sort(consts->begin(), consts->end(),
                  [&](int32_t a, int32_t b) { return c.nodes[(size_t)a].pos < c.nodes[(size_t)b].pos; });
    }

    // Emits one job over the constants group[0..ng).
    bool job(int32_t prog, const ProgMeta& m, const int32_t* group, int ng, bool first, std::string* err) {
        const size_t nn = c.nodes.size();
        act.assign(nn, 0);
        kidx.assign(nn, -1);
        for (int k = 0; k < ng; ++k) { act[(size_t)group[k]] = 1; kidx[(size_t)group[k]] = k; }
        // Nodes are stored children-first, so one forward sweep propagates activity up.
        for (size_t v = 0; v < nn; ++v) {
            const Node& n = c.nodes[v];
            if (n.op < 0 || !live[v]) continue;
            for (int j = 0; j < arity(n.op); ++j) {
                float dummy;
                if (act[(size_t)n.ch[j]] && edge_kind((int32_t)v, j, &dummy) != 0) act[v] = 1;
            }
        }
        need.assign(nn, 0);
        for (size_t v = 0; v < nn; ++v) {
            const Node& n = c.nodes[v];
            if (!act[v] || n.op < 0) continue;
            std::vector<int32_t> kids = active_kids((int32_t)v);
            int nd = 1;
            for (size_t j = 0; j < kids.size(); ++j) {
                int cn = need[(size_t)kids[j]];
                nd = std::max(nd, j + 1 < kids.size() ? 1 + cn : cn);
            }
            need[v] = nd;
        }
        AdjJob J{};
        J.code_off = m.code_off;
        J.ncode = m.ncode;
        J.prog = prog;
        J.const_base = (uint32_t)out->consts.size();
        J.nconst = (uint32_t)ng;
        J.flags = (m.flags & (kEmpty | kInvalid)) | (first ? kAdjFitness : 0u);
        const size_t rev0 = out->rev.size();
        std::vector<int32_t> seg_of;     // per reverse instruction
        std::vector<uint32_t> e_ins;     // per stored partial: instruction (relative)
        std::vector<uint8_t> e_hi;       // per stored partial: 1 = other-operand side
        int32_t cur_seg = 0;
        int max_slot = 0;
        auto put_rev = [&](uint32_t kind, int src, int dst, int kindp, float v, int32_t e) {
            RevIns r{};
            r.w = kind | ((uint32_t)src << 4) | ((uint32_t)dst << 8);
            if (kindp == 2) {
                r.w |= kRevStored | ((uint32_t)e << 16);
                cur_seg = e / B;
            } else {
                std::memcpy(&r.p, &v, 4);
            }
            out->rev.push_back(r);
            seg_of.push_back(cur_seg);
        };
        if (ng > 0) {
            const int32_t root = c.root_node;
            if (c.nodes[(size_t)root].op < 0) {
                put_rev(kRevGrad, 0, kidx[(size_t)root], 1, 1.0f, 0);
            } else {
                if (need[(size_t)root] > kAdjSlots) {
                    if (err) *err = "program " + std::to_string(prog) + " needs " + std::to_string(need[(size_t)root]) +
                                    " adjoint slots (max " + std::to_string(kAdjSlots) + ")";
                    return false;
                }
                // Preorder walk with an explicit stack of (node, slot, next child).
                struct Fr { int32_t v; int slot; std::vector<int32_t> kids; size_t j; };
                std::vector<Fr> st;
                st.push_back(Fr{root, 0, active_kids(root), 0});
                while (!st.empty()) {
                    Fr& f = st.back();
                    if (f.j == f.kids.size()) { st.pop_back(); continue; }
                    const int32_t v = f.v, ch = f.kids[f.j];
                    const bool last = f.j + 1 == f.kids.size();
                    const int slot = f.slot;
                    ++f.j;
                    const Node& n = c.nodes[(size_t)v];
                    int jj = 0;
                    while (n.ch[jj] != ch) ++jj;
                    float sv = 0.0f;
                    const int kind = edge_kind(v, jj, &sv);
                    int32_t e = -1;
                    if (kind == 2) {
                        e = (int32_t)e_ins.size();
                        const uint32_t rel = (uint32_t)(c.ins_at[(size_t)v] - (int32_t)m.code_off);
                        e_ins.push_back(rel);
                        const bool hi = arity(n.op) == 3 ? (jj == 2) : (c.side[(size_t)ch] == 1);
                        e_hi.push_back(hi ? 1 : 0);
                    }
                    if (c.nodes[(size_t)ch].op < 0) {
                        put_rev(kRevGrad, slot, kidx[(size_t)ch], kind, sv, e);
                    } else {
                        const int cs = last ? slot : slot + 1;
                        max_slot = std::max(max_slot, cs + 1);
                        put_rev(kRevAdj, slot, cs, kind, sv, e);
                        st.push_back(Fr{ch, cs, active_kids(ch), 0});
                    }
                }
            }
        }
        if (e_ins.size() >= 0xffffu) {
            if (err) *err = "program " + std::to_string(prog) + ": too many stored partials";
            return false;
        }
        const int32_t nst = (int32_t)e_ins.size();
        const int32_t nseg = std::max<int32_t>(1, (nst + B - 1) / B);
        J.seg_off = (uint32_t)out->segs.size();
        J.nseg = (uint32_t)nseg;
        size_t r = rev0;
        for (int32_t s = 0; s < nseg; ++s) {
            AdjSeg sg{};
            sg.rev_begin = (uint32_t)r;
            while (r < out->rev.size() && seg_of[r - rev0] == s) ++r;
            sg.rev_end = (uint32_t)r;
            // The segment's stores, one per instruction, in instruction order.
            std::vector<AdjStore> ss;
            for (int32_t e = s * B; e < std::min(nst, (s + 1) * B); ++e)
                ss.push_back(AdjStore{e_ins[(size_t)e], (uint32_t)(e - s * B + 1) << (e_hi[(size_t)e] ? 16 : 0)});
            std::sort(ss.begin(), ss.end(), [](const AdjStore& x, const AdjStore& y) { return x.pc < y.pc; });
            sg.st_begin = (uint32_t)out->stores.size();
            for (const AdjStore& q : ss) {
                if (out->stores.size() > sg.st_begin && out->stores.back().pc == q.pc) out->stores.back().slots |= q.slots;
                else out->stores.push_back(q);
            }
            sg.st_end = (uint32_t)out->stores.size();
            if (s == 0) sg.fwd_end = m.ncode;
            else sg.fwd_end = sg.st_end > sg.st_begin ? out->stores.back().pc + 1 : 0;
            out->segs.push_back(sg);
        }
        J.cost = m.ncode * (uint32_t)nseg + (uint32_t)(out->rev.size() - rev0);
        J.cols = 0;
        for (uint32_t q = m.code_off; q < m.code_off + m.ncode; ++q) {
            const uint32_t h = t.code[q].code & 0xffu, col = t.code[q].arg;
            if ((h == H_LOAD_A || h == H_PUSH_A || (h >= H_BIN_A && h < H_BIN_S)) && col < 32) J.cols |= 1u << col;
        }
        for (int k = 0; k < ng; ++k) {
            AdjConst ac{};
            ac.prog = prog;
            ac.pos = c.nodes[(size_t)group[k]].pos;
            ac.ins = 0xffffffffu;
            for (uint32_t q = m.code_off; q < m.code_off + m.ncode; ++q)
                if (t.const_pos[q] == (uint32_t)ac.pos + 1) { ac.ins = q; break; }
            out->consts.push_back(ac);
        }
        out->jobs.push_back(J);
        out->max_seg = std::max(out->max_seg, nseg);
        out->max_adj_slots = std::max(out->max_adj_slots, max_slot);
        out->stored_partials += nst;
        out->total_rev += (int64_t)(out->rev.size() - rev0);
        return true;
    }

    // Active children of v, ordered so that the one needing most adjoint slots
    // comes last (it reuses v's slot).
    std::vector<int32_t> active_kids(int32_t v) const {
        const Node& n = c.nodes[(size_t)v];
        std::vector<int32_t> k;
        for (int j = 0; j < arity(n.op); ++j) {
            float dummy;
            if (act[(size_t)n.ch[j]] && edge_kind(v, j, &dummy) != 0) k.push_back(n.ch[j]);
        }
        std::stable_sort(k.begin(), k.end(), [&](int32_t a, int32_t b) {
            return need[(size_t)a] < need[(size_t)b];
        });
        return k;
    }
};

}  // namespace

bool compile_adjoint(const SymbolTable& st, const int32_t* phenotype, const float* ephemeral, const int32_t* size,
                     int32_t stride, int32_t n, int32_t ncol, int32_t B, AdjTape* out, std::string* err) {
    if (B <= 0 || (B & (B - 1))) {
        if (err) *err = "partial store size must be a power of two";
        return false;
    }
    *out = AdjTape{};
    Tape& t = out->fwd;
    Compiler c(st, ncol, true);
    c.track = true;
    AdjBuilder ab(c, t, B, out);
    out->prog_jobs.push_back(0);
    out->prog_consts.push_back(0);
    std::vector<int32_t> consts;
    for (int32_t i = 0; i < n; ++i) {
        ProgMeta m{};
        m.code_off = (uint32_t)t.code.size();
        m.size = size[i];
        const int32_t sz = size[i];
        if (sz < 0 || sz > stride) {
            if (err) *err = "program " + std::to_string(i) + ": size " + std::to_string(sz) + " outside [0, " +
                            std::to_string(stride) + "]";
            return false;
        }
        bool compiled = false;
        if (sz == 0) {
            m.flags = kEmpty;
            t.stats.n_empty++;
        } else {
            int slots = c.compile(phenotype + (size_t)i * stride, ephemeral + (size_t)i * stride, sz, &t);
            if (slots < 0) {
                t.code.resize(m.code_off);
                t.const_pos.resize(m.code_off);
                t.code.push_back(Ins{H_LOAD_C, 0x7fc00000u});
                t.const_pos.push_back(0);
                m.flags = kInvalid;
                t.stats.n_invalid++;
                slots = 0;
            } else {
                compiled = true;
            }
            if (slots > kMaxSlots) {
                if (err) *err = "program " + std::to_string(i) + " needs " + std::to_string(slots) +
                                " register slots (max " + std::to_string(kMaxSlots) + ")";
                return false;
            }
            m.flags |= (uint32_t)slots << 8;
            t.stats.total_size += sz;
            t.stats.max_slots = std::max(t.stats.max_slots, slots);
        }
        m.ncode = (uint32_t)t.code.size() - m.code_off;
        t.code.push_back(Ins{H_END, m.flags & (kEmpty | kInvalid)});
        t.const_pos.push_back(0);
        t.stats.total_code += m.ncode;
        t.stats.max_code = std::max<int32_t>(t.stats.max_code, (int32_t)m.ncode);
        t.meta.push_back(m);
        consts.clear();
        if (compiled) ab.live_consts(&consts);
        if (consts.empty()) {
            if (!ab.job(i, m, nullptr, 0, true, err)) return false;
        } else {
            for (size_t g = 0; g < consts.size(); g += kAdjMaxConsts) {
                int ng = (int)std::min<size_t>(kAdjMaxConsts, consts.size() - g);
                if (!ab.job(i, m, consts.data() + g, ng, g == 0, err)) return false;
            }
        }
        out->prog_jobs.push_back((uint32_t)out->jobs.size());
        out->prog_consts.push_back((uint32_t)out->consts.size());
    }
    for (size_t k = 0; k < t.code.size(); ++k) {
        uint32_t h = t.code[k].code & 0xffu;
        if (h == H_LOAD_A || h == H_PUSH_A || (h >= H_BIN_A && h < H_BIN_S))
            t.stats.max_col = std::max<int32_t>(t.stats.max_col, (int32_t)t.code[k].arg);
    }
    return true;
}

}  // namespace ppi
