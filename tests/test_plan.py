"""CPU: join plan documents (the query_tree format of the reference's
include/execution.hpp, selected by IDPHandler::SetQueryTree,
src/pipeline.cc:251-279) are read, validated and rejected with SPJ_EINVAL
when the GPU phase sequence cannot run them."""
import json

import pytest


def send(side):
    return {"operator": "HOST_FUNC_SEND_DATA_OPT", "buffer_name": side}


def part(side, op="COMPOUND_FUNC_GLB_PARTITION"):
    pre = "Build" if side == "left" else "Probe"
    return {"operator": op, "inputs": ["unused", side], "outputs": [pre + "_arr", pre + "_hist"],
            "packet_size": 8, "partition_type": 0 if side == "left" else 1}


SYNC = {"operator": "CONTROL_FUNC_SYNC_THREADS"}
JOIN = {"operator": "COMPOUND_FUNC_RNS_JOIN", "inputs": ["Build_arr", "Build_hist", "Probe_arr", "Probe_hist"],
        "outputs": ["join_result"], "packet_size": 8, "join_type": "", "comment": "first join"}
RECV = {"operator": "HOST_FUNC_RECV_DATA_OPT", "node_name": "join_result"}


def plain_tree():
    return [send("left"), send("right"), SYNC, part("left"), part("right"), JOIN, RECV]


def bg_tree():
    return [send("left"), send("right"), SYNC,
            {"operator": "COMPOUND_FUNC_BG_BROADCAST", "inputs": ["unused", "left"],
             "outputs": ["Build_arr", "build_side_r2r_xfer_cnt"]},
            {"operator": "HOST_FUNC_CHIPWISE_BROADCAST", "inputs": ["Build_arr", "build_side_r2r_xfer_cnt"],
             "outputs": ["Replicated_R_Tuples"]},
            part("right", "COMPOUND_FUNC_GLB_CHIPWISE_PARTITION"),
            {"operator": "HOST_FUNC_ROTATE_AND_STREAM", "input": "Probe_arr", "output": "S_rnc_join_key",
             "histogram": "Probe_hist", "packet_size": 8, "rankgroup_aware": 0},
            {"operator": "COMPOUND_FUNC_IDP_JOIN", "inputs": ["Replicated_R_Tuples", "S_rnc_join_key"],
             "outputs": ["join_result"], "packet_size": 8},
            RECV]


def doc(tree):
    return json.dumps({"query_name": "", "tables": [], "query_tree": tree, "query_string": []}, indent=4)


def test_plain_and_replicating_plans(sj):
    assert sj.plan_check(doc(plain_tree())) == (7, False)
    assert sj.plan_check(doc(bg_tree())) == (9, True)
    # the reference's documents are C string macros with line-continuation
    # padding: runs of spaces between tokens are whitespace
    padded = doc(plain_tree()).replace("\n", "  \t \r\n")
    assert sj.plan_check(padded) == (7, False)


def test_json_reader_details(sj):
    tree = plain_tree()
    tree[5] = dict(JOIN, comment="café \"q\" \\ / \t", packet_size=8.0e0, extra=[None, True, False, -1.5e-3, {}])
    text = doc(tree).replace("caf\\u00e9", "caf\\u00E9")
    assert sj.plan_check(text) == (7, False)
    # \u escapes in an operator name are decoded before the name is matched
    esc = doc(plain_tree()).replace('"CONTROL_FUNC_SYNC_THREADS"', '"CONTROL_FUNC_SYNC_\\u0054HREADS"')
    assert sj.plan_check(esc) == (7, False)


@pytest.mark.parametrize("text", [
    "", "{", "[]", '{"query_tree": []}', '{"query_tree": {}}', '{"query_tree": [1]}',
    '{"query_tree": [{"operator": 5}]}', '{"query_tree": [{}]}', "{} x", '{"a": tru}', '{"a": "\\x"}',
    '{"a": 01e}', "[" * 100 + "]" * 100,
])
def test_malformed_documents_rejected(sj, text):
    with pytest.raises(sj.SpjError) as e:
        sj.plan_check(text)
    assert e.value.code == -1


def mutate(tree, i, node=None, drop=False):
    t = list(tree)
    if drop:
        del t[i]
    else:
        t[i] = node
    return t


@pytest.mark.parametrize("tree,why", [
    (mutate(plain_tree(), 5, {"operator": "DPU_FUNC_NESTED_LOOP_JOIN"}), "not one of a join plan's operators"),
    (mutate(plain_tree(), 2, {"operator": "NO_SUCH_OPERATOR"}), "not one of a join plan's operators"),
    ([send("left"), part("left"), part("right"), send("right"), JOIN, RECV], "before its relation is sent"),
    ([send("left"), send("right"), part("left"), JOIN, part("right"), RECV], "before both relations"),
    (plain_tree()[:6] + [JOIN, RECV], "one join"),
    (plain_tree()[:6], "does not receive"),
    (mutate(plain_tree(), 6, drop=True)[:5] + [RECV, JOIN], "before the join"),
    (plain_tree() + [RECV], "received once"),
    ([send("left"), send("left")] + plain_tree()[2:], "sent twice"),
    (mutate(plain_tree(), 0, {"operator": "HOST_FUNC_SEND_DATA_OPT", "buffer_name": "middle"}), "left or right"),
    (mutate(plain_tree(), 4, part("left")), "partitioned twice"),
    (mutate(plain_tree(), 3, {"operator": "COMPOUND_FUNC_GLB_PARTITION", "inputs": ["left"]}), "needs inputs"),
    (mutate(bg_tree(), 3, {"operator": "COMPOUND_FUNC_BG_BROADCAST", "inputs": ["unused", "right"]}),
     "replicates the left"),
    (mutate(bg_tree(), 5, part("left", "COMPOUND_FUNC_GLB_CHIPWISE_PARTITION")), "partitions the right"),
    (bg_tree()[:3] + [bg_tree()[4], bg_tree()[3]] + bg_tree()[5:], "before BG_BROADCAST"),
    ([send("left"), send("right"), bg_tree()[6], part("left"), part("right"), JOIN, RECV],
     "before the right partition"),
    (plain_tree()[:5] + [RECV], "before the join"),
])
def test_invalid_plans_rejected(sj, tree, why):
    with pytest.raises(sj.SpjError) as e:
        sj.plan_check(doc(tree))
    assert e.value.code == -1
    assert why in str(e.value), str(e.value)


def header_plans():
    """The join_json / bg_join_json macros of include/execution.hpp, as text."""
    import os
    import re
    src = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                            "execution.hpp")).read()
    out = {}
    for name in ("join_json", "bg_join_json"):
        body = re.search(r"#define %s \\\n((?:    \".*\"(?: \\)?\n)+)" % name, src).group(1)
        pieces = re.findall(r'"((?:[^"\\]|\\.)*)"', body)
        out[name] = "".join(p.replace('\\"', '"').replace("\\\\", "\\") for p in pieces)
    return out


def test_execution_header_plans(sj):
    """include/execution.hpp's two plan strings are valid plans of the shape
    StartJoin picks between."""
    p = header_plans()
    assert sj.plan_check(p["join_json"]) == (7, False)
    assert sj.plan_check(p["bg_join_json"]) == (9, True)
    assert json.loads(p["join_json"])["query_tree"] == plain_tree()
    assert json.loads(p["bg_join_json"])["query_tree"] == bg_tree()
