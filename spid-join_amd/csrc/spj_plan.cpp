// spj_plan.cpp -- join plan documents and the per-operator timeline.
//
// A small JSON reader (objects, arrays, strings, numbers, literals) takes the
// plan document; the operators of the two join plans are mapped onto the GPU
// phases and checked for an order the phase sequence can run: each relation
// sent, then partitioned once, one join after both partitions, the result
// received after the join.
#include "spj_plan.hpp"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "spidjoin.h"
#include "spj_capi_internal.hpp"

namespace spj {

namespace {

struct JVal {
    enum Type { Null, Bool, Num, Str, Arr, Obj } t = Null;
    bool b = false;
    double n = 0;
    std::string s;
    std::vector<JVal> a;
    std::vector<std::pair<std::string, JVal>> o;
    const JVal *get(const char *k) const
    {
        for (const auto &kv : o)
            if (kv.first == k) return &kv.second;
        return nullptr;
    }
};

class JReader {
  public:
    explicit JReader(const char *text) : p_(text), end_(text + strlen(text)) {}
    bool document(JVal &v)
    {
        if (!value(v, 0)) return false;
        ws();
        if (p_ != end_) return error("trailing characters");
        return true;
    }
    const std::string &message() const { return err_; }
    size_t offset(const char *base) const { return (size_t)(p_ - base); }

  private:
    const char *p_, *end_;
    std::string err_;

    bool error(const char *what)
    {
        if (err_.empty()) err_ = what;
        return false;
    }
    void ws()
    {
        while (p_ < end_ && (*p_ == ' ' || *p_ == '\t' || *p_ == '\n' || *p_ == '\r')) p_++;
    }
    bool literal(const char *word)
    {
        const size_t n = strlen(word);
        if ((size_t)(end_ - p_) < n || strncmp(p_, word, n) != 0) return error("unknown literal");
        p_ += n;
        return true;
    }
    static void utf8(std::string &s, uint32_t c)
    {
        if (c < 0x80) {
            s += (char)c;
        } else if (c < 0x800) {
            s += (char)(0xC0 | (c >> 6));
            s += (char)(0x80 | (c & 0x3F));
        } else if (c < 0x10000) {
            s += (char)(0xE0 | (c >> 12));
            s += (char)(0x80 | ((c >> 6) & 0x3F));
            s += (char)(0x80 | (c & 0x3F));
        } else {
            s += (char)(0xF0 | (c >> 18));
            s += (char)(0x80 | ((c >> 12) & 0x3F));
            s += (char)(0x80 | ((c >> 6) & 0x3F));
            s += (char)(0x80 | (c & 0x3F));
        }
    }
    bool hex4(uint32_t &c)
    {
        if (end_ - p_ < 4) return error("short \\u escape");
        c = 0;
        for (int i = 0; i < 4; i++) {
            const char h = *p_++;
            c <<= 4;
            if (h >= '0' && h <= '9') c |= (uint32_t)(h - '0');
            else if (h >= 'a' && h <= 'f') c |= (uint32_t)(h - 'a' + 10);
            else if (h >= 'A' && h <= 'F') c |= (uint32_t)(h - 'A' + 10);
            else return error("bad \\u escape");
        }
        return true;
    }
    bool string(std::string &s)
    {
        if (p_ >= end_ || *p_ != '"') return error("expected a string");
        p_++;
        while (p_ < end_ && *p_ != '"') {
            const char ch = *p_++;
            if ((unsigned char)ch < 0x20) return error("control character in a string");
            if (ch != '\\') {
                s += ch;
                continue;
            }
            if (p_ >= end_) return error("unterminated escape");
            const char e = *p_++;
            switch (e) {
            case '"': s += '"'; break;
            case '\\': s += '\\'; break;
            case '/': s += '/'; break;
            case 'b': s += '\b'; break;
            case 'f': s += '\f'; break;
            case 'n': s += '\n'; break;
            case 'r': s += '\r'; break;
            case 't': s += '\t'; break;
            case 'u': {
                uint32_t c;
                if (!hex4(c)) return false;
                if (c >= 0xD800 && c < 0xDC00 && end_ - p_ >= 6 && p_[0] == '\\' && p_[1] == 'u') {
                    const char *save = p_;
                    p_ += 2;
                    uint32_t lo;
                    if (!hex4(lo)) return false;
                    if (lo >= 0xDC00 && lo < 0xE000) c = 0x10000 + ((c - 0xD800) << 10) + (lo - 0xDC00);
                    else p_ = save;
                }
                utf8(s, c);
                break;
            }
            default: return error("bad escape");
            }
        }
        if (p_ >= end_) return error("unterminated string");
        p_++;
        return true;
    }
    bool number(double &v)
    {
        const char *b = p_;
        if (p_ < end_ && *p_ == '-') p_++;
        if (p_ >= end_ || *p_ < '0' || *p_ > '9') return error("bad number");
        while (p_ < end_ && ((*p_ >= '0' && *p_ <= '9') || *p_ == '.' || *p_ == 'e' || *p_ == 'E' || *p_ == '+' ||
                             *p_ == '-'))
            p_++;
        const std::string t(b, p_);
        char *e = nullptr;
        v = strtod(t.c_str(), &e);
        if (!e || *e != '\0') return error("bad number");
        return true;
    }
    bool value(JVal &v, int depth)
    {
        if (depth > 64) return error("nesting too deep");
        ws();
        if (p_ >= end_) return error("unexpected end");
        const char c = *p_;
        if (c == '{') {
            v.t = JVal::Obj;
            p_++;
            ws();
            if (p_ < end_ && *p_ == '}') {
                p_++;
                return true;
            }
            for (;;) {
                ws();
                std::pair<std::string, JVal> kv;
                if (!string(kv.first)) return false;
                ws();
                if (p_ >= end_ || *p_ != ':') return error("expected ':'");
                p_++;
                if (!value(kv.second, depth + 1)) return false;
                v.o.push_back(std::move(kv));
                ws();
                if (p_ < end_ && *p_ == ',') {
                    p_++;
                    continue;
                }
                if (p_ < end_ && *p_ == '}') {
                    p_++;
                    return true;
                }
                return error("expected ',' or '}'");
            }
        }
        if (c == '[') {
            v.t = JVal::Arr;
            p_++;
            ws();
            if (p_ < end_ && *p_ == ']') {
                p_++;
                return true;
            }
            for (;;) {
                v.a.emplace_back();
                if (!value(v.a.back(), depth + 1)) return false;
                ws();
                if (p_ < end_ && *p_ == ',') {
                    p_++;
                    continue;
                }
                if (p_ < end_ && *p_ == ']') {
                    p_++;
                    return true;
                }
                return error("expected ',' or ']'");
            }
        }
        if (c == '"') {
            v.t = JVal::Str;
            return string(v.s);
        }
        if (c == 't') {
            v.t = JVal::Bool;
            v.b = true;
            return literal("true");
        }
        if (c == 'f') {
            v.t = JVal::Bool;
            return literal("false");
        }
        if (c == 'n') return literal("null");
        v.t = JVal::Num;
        return number(v.n);
    }
};

struct OpName {
    const char *name;
    PlanOp op;
};
const OpName kOps[] = {
    {"HOST_FUNC_SEND_DATA_OPT", PlanOp::Send},
    {"CONTROL_FUNC_SYNC_THREADS", PlanOp::Sync},
    {"COMPOUND_FUNC_GLB_PARTITION", PlanOp::Partition},
    {"COMPOUND_FUNC_GLB_CHIPWISE_PARTITION", PlanOp::ChipwisePartition},
    {"COMPOUND_FUNC_BG_BROADCAST", PlanOp::BgBroadcast},
    {"HOST_FUNC_CHIPWISE_BROADCAST", PlanOp::ChipwiseBroadcast},
    {"HOST_FUNC_ROTATE_AND_STREAM", PlanOp::RotateAndStream},
    {"COMPOUND_FUNC_RNS_JOIN", PlanOp::RnsJoin},
    {"COMPOUND_FUNC_IDP_JOIN", PlanOp::IdpJoin},
    {"HOST_FUNC_RECV_DATA_OPT", PlanOp::Recv},
};

int side_of(const char *s)
{
    if (strcmp(s, "left") == 0) return 0;
    if (strcmp(s, "right") == 0) return 1;
    return -1;
}

// inputs as the timeline prints them (join_internals.hpp:63-83): the
// "input" string, else the "inputs" strings joined by ','
std::string inputs_of(const JVal &node)
{
    if (const JVal *in = node.get("input")) return in->t == JVal::Str ? in->s : std::string();
    std::string r;
    if (const JVal *ins = node.get("inputs")) {
        for (size_t i = 0; i < ins->a.size(); i++) {
            if (i) r += ",";
            if (ins->a[i].t == JVal::Str) r += ins->a[i].s;
        }
    }
    return r;
}

// the relation a partitioning operator works on: inputs[1]
int partition_side(const JVal &node)
{
    const JVal *ins = node.get("inputs");
    if (!ins || ins->t != JVal::Arr || ins->a.size() < 2 || ins->a[1].t != JVal::Str) return -1;
    return side_of(ins->a[1].s.c_str());
}

PlanNode make(PlanOp op, const char *inputs, int side)
{
    PlanNode n;
    n.op = op;
    for (const OpName &o : kOps)
        if (o.op == op) n.name = o.name;
    n.inputs = inputs;
    n.side = side;
    return n;
}

}  // namespace

int parse_plan(const char *json, Plan *out)
{
    if (!json || !out) return fail(SPJ_EINVAL, "null plan");
    JVal root;
    JReader rd(json);
    if (!rd.document(root)) return fail(SPJ_EINVAL, "plan is not JSON: %s at byte %zu", rd.message().c_str(),
                                        rd.offset(json));
    const JVal *tree = root.t == JVal::Obj ? root.get("query_tree") : nullptr;
    if (!tree || tree->t != JVal::Arr || tree->a.empty()) return fail(SPJ_EINVAL, "plan has no query_tree array");
    Plan plan;
    bool sent[2] = {false, false}, parted[2] = {false, false};
    int joins = 0, recvs = 0;
    bool broadcast = false;
    for (size_t i = 0; i < tree->a.size(); i++) {
        const JVal &nd = tree->a[i];
        const JVal *op = nd.t == JVal::Obj ? nd.get("operator") : nullptr;
        if (!op || op->t != JVal::Str) return fail(SPJ_EINVAL, "plan node %zu has no operator", i);
        const OpName *on = nullptr;
        for (const OpName &o : kOps)
            if (op->s == o.name) on = &o;
        if (!on) return fail(SPJ_EINVAL, "plan node %zu: operator %s is not one of a join plan's operators", i,
                             op->s.c_str());
        PlanNode n;
        n.op = on->op;
        n.name = on->name;
        n.inputs = inputs_of(nd);
        switch (n.op) {
        case PlanOp::Send: {
            const JVal *b = nd.get("buffer_name");
            n.side = (b && b->t == JVal::Str) ? side_of(b->s.c_str()) : -1;
            if (n.side < 0) return fail(SPJ_EINVAL, "plan node %zu: SEND_DATA_OPT buffer_name must be left or right", i);
            if (sent[n.side]) return fail(SPJ_EINVAL, "plan node %zu: %s sent twice", i, n.side ? "right" : "left");
            sent[n.side] = true;
            break;
        }
        case PlanOp::Partition:
        case PlanOp::ChipwisePartition:
        case PlanOp::BgBroadcast: {
            n.side = partition_side(nd);
            if (n.side < 0) return fail(SPJ_EINVAL, "plan node %zu: %s needs inputs [_, left|right]", i, on->name);
            if (n.op == PlanOp::ChipwisePartition && n.side != 1)
                return fail(SPJ_EINVAL, "plan node %zu: GLB_CHIPWISE_PARTITION partitions the right relation", i);
            if (n.op == PlanOp::BgBroadcast && n.side != 0)
                return fail(SPJ_EINVAL, "plan node %zu: BG_BROADCAST replicates the left relation", i);
            if (!sent[n.side]) return fail(SPJ_EINVAL, "plan node %zu: %s before its relation is sent", i, on->name);
            if (parted[n.side])
                return fail(SPJ_EINVAL, "plan node %zu: %s partitioned twice", i, n.side ? "right" : "left");
            parted[n.side] = true;
            if (n.op == PlanOp::BgBroadcast) {
                broadcast = true;
                plan.replicate = true;
            }
            break;
        }
        case PlanOp::ChipwiseBroadcast:
            if (!broadcast) return fail(SPJ_EINVAL, "plan node %zu: CHIPWISE_BROADCAST before BG_BROADCAST", i);
            break;
        case PlanOp::RotateAndStream:
            if (!parted[1]) return fail(SPJ_EINVAL, "plan node %zu: ROTATE_AND_STREAM before the right partition", i);
            break;
        case PlanOp::RnsJoin:
        case PlanOp::IdpJoin:
            if (!parted[0] || !parted[1])
                return fail(SPJ_EINVAL, "plan node %zu: %s before both relations are partitioned", i, on->name);
            if (++joins > 1) return fail(SPJ_EINVAL, "plan node %zu: a plan has one join", i);
            break;
        case PlanOp::Recv:
            if (!joins) return fail(SPJ_EINVAL, "plan node %zu: RECV_DATA_OPT before the join", i);
            if (++recvs > 1) return fail(SPJ_EINVAL, "plan node %zu: the result is received once", i);
            break;
        case PlanOp::Sync: break;
        }
        plan.nodes.push_back(std::move(n));
    }
    if (!joins) return fail(SPJ_EINVAL, "plan has no join operator");
    if (!recvs) return fail(SPJ_EINVAL, "plan does not receive the join result");
    *out = std::move(plan);
    return SPJ_OK;
}

Plan default_plan(bool replicate)
{
    Plan p;
    p.replicate = replicate;
    p.nodes.push_back(make(PlanOp::Send, "", 0));
    p.nodes.push_back(make(PlanOp::Send, "", 1));
    p.nodes.push_back(make(PlanOp::Sync, "", -1));
    if (!replicate) {
        p.nodes.push_back(make(PlanOp::Partition, "unused,left", 0));
        p.nodes.push_back(make(PlanOp::Partition, "unused,right", 1));
        p.nodes.push_back(make(PlanOp::RnsJoin, "Build_arr,Build_hist,Probe_arr,Probe_hist", -1));
    } else {
        p.nodes.push_back(make(PlanOp::BgBroadcast, "unused,left", 0));
        p.nodes.push_back(make(PlanOp::ChipwiseBroadcast, "Build_arr,build_side_r2r_xfer_cnt", -1));
        p.nodes.push_back(make(PlanOp::ChipwisePartition, "unused,right", 1));
        p.nodes.push_back(make(PlanOp::RotateAndStream, "Probe_arr", -1));
        p.nodes.push_back(make(PlanOp::IdpJoin, "Replicated_R_Tuples,S_rnc_join_key", -1));
    }
    p.nodes.push_back(make(PlanOp::Recv, "", -1));
    return p;
}

std::string timeline_text(const Plan &plan, const std::vector<NodeTime> &t, int64_t t0_ns, int64_t start_ns,
                          int rank_group, const std::vector<uint64_t> &parted)
{
    std::string s;
    char line[512];
    const double off = (double)(t0_ns - start_ns) / 1e6;
    for (size_t i = 0; i < plan.nodes.size() && i < t.size(); i++) {
        // single-precision milliseconds printed with %f, as the reference's
        // std::to_string(float)
        const float a = (float)(off + t[i].start_ms), b = (float)(off + t[i].stop_ms);
        snprintf(line, sizeof line, "%d-0-%zu-%f-%f-%s-%s\n", rank_group, i, (double)a, (double)b,
                 plan.nodes[i].name.c_str(), plan.nodes[i].inputs.c_str());
        s += line;
    }
    for (uint64_t n : parted) {
        snprintf(line, sizeof line, "rnc_xfer_bytes %llu rnc_num_packets %llu\n", (unsigned long long)(n * 8),
                 (unsigned long long)n);
        s += line;
    }
    return s;
}

}  // namespace spj
