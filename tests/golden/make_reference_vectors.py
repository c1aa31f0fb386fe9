"""Writes tests/golden/reference_vectors.json from the reference's own tests.

Run in a container where the reference tree is mounted (it reads the test
files as text; nothing of the reference is executed):

    python tests/golden/make_reference_vectors.py /root/reference

Each vector keeps its provenance (file and line range).  The JSON holds only
data: the matrices and index lists the reference tests assert against, and the
parameter values of the reference's JSON input files.
"""
import json
import re
import sys
from pathlib import Path


def matrix(lines, start, end):
    """Numbers of the literal on lines start..end (1-based, inclusive), one
    row per text line that holds numbers."""
    rows = []
    for ln in lines[start - 1:end]:
        body = ln.split("=", 1)[1] if ("=" in ln and "[" in ln.split("=", 1)[1]) else ln
        body = body.replace("1.0e-03 *", "")
        nums = re.findall(r"-?\d+\.?\d*(?:[eE][-+]?\d+)?", body.replace("[", " ").replace("]", " "))
        if nums:
            rows.append([float(x) for x in nums])
    return rows


def index_list(line):
    m = re.search(r"\[\[([\d,\s]+)\]\]", line)
    return [int(x) for x in m.group(1).split(",")]


_EPS = 2.220446049250313e-16


def _vec3(item):
    """One 3-vector literal of MiscTest.jl's minimumDistance cases; only the
    forms that file uses are accepted (nothing is evaluated)."""
    item = item.strip()
    if item == "zeros(3)":
        return [0.0, 0.0, 0.0]
    if item == "ones(3)":
        return [1.0, 1.0, 1.0]
    m = re.fullmatch(r"(-?\d+\.?\d*) \* ones\(3\)", item)
    if m:
        return [float(m.group(1))] * 3
    if item == "ones(3) .+ eps(Float64)^2":
        return [1.0 + _EPS ** 2] * 3
    m = re.fullmatch(r"\[(-?[\d.]+), (-?[\d.]+), (-?[\d.]+)\]", item)
    if m:
        return [float(m.group(k)) for k in (1, 2, 3)]
    raise ValueError(f"unexpected vector literal {item!r}")


def _split_top(text):
    out, depth, cur = [], 0, ""
    for ch in text:
        if ch == "[" or ch == "(":
            depth += 1
        elif ch == "]" or ch == ")":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur)
            cur = ""
        else:
            cur += ch
    out.append(cur)
    return out


def min_distance_cases(lines, a, b):
    """minimumDistance cases of MiscTest.jl: inputs and expected
    (distSq, dDistSqDt, L1, L2), with assignments carried over between cases
    as the test file does."""
    env, cases = {}, []
    for ln in lines[a - 1:b]:
        m = re.match(r"\s*(\w+), (\w+) = (.*)$", ln)
        if m and m.group(1) in ("x0", "x1", "y0", "y1", "vx0", "vx1", "vy0", "vy1"):
            vals = _split_top(m.group(3))
            env[m.group(1)], env[m.group(2)] = _vec3(vals[0]), _vec3(vals[1])
            continue
        m = re.match(r"\s*@test \(distSq, dDistSqDt, L1, L2\) == \((.*)\)\s*$", ln)
        if m:
            exp = [float(v) for v in m.group(1).split(",")]
            cases.append({"in": {k: list(env[k]) for k in
                                 ("x0", "x1", "y0", "y1", "vx0", "vx1", "vy0", "vy1")},
                          "expect": exp})
    return cases


def _jl_matrix(text):
    """Julia matrix literal '[a b c; d e f]' (optionally typed) -> rows."""
    body = re.sub(r"^\w*\[|\]$", "", text.strip())
    rows = [r.split() for r in body.split(";")]
    return [[float(v) for v in r] for r in rows]


def collision_cases(lines, a, b):
    """detectCollision networks and expected results of TopologyTest.jl, in
    the order they appear (calls may span lines)."""
    text = "\n".join(lines[a - 1:b])
    events = []
    for m in re.finditer(r"(links|coord|label|nodeVel|bVec) = ([\w]*\[[^\]]*\])", text):
        events.append((m.start(), "field", m.group(1), _jl_matrix(m.group(2))))
    for m in re.finditer(r"detectCollision\(\s*dlnParams,\s*network,\s*"
                         r"(skipSegs|\[[^\]]*\])\s*,?\s*\)", text):
        call = m.group(1)
        skip = [] if call == "skipSegs" else [
            [int(x), int(y)] for x, y in re.findall(r"\((\d+), (\d+)\)", call)]
        events.append((m.start(), "call", skip, None))
    for m in re.finditer(r"\((true|false), :(\w+), ([-\d., ]+)\)", text):
        exp = [m.group(1) == "true", m.group(2)] + [float(v) for v in m.group(3).split(",")]
        events.append((m.start(), "expect", exp, None))
    nets, checks, cur = [], [], None
    for _, kind, x, y in sorted(events, key=lambda e: e[0]):
        if kind == "field":
            if x == "links":
                cur = {}
                nets.append(cur)
            cur[x] = y
        elif kind == "call":
            checks.append({"net": len(nets) - 1, "skip": x, "expect": None})
        else:
            for c in checks:
                if c["expect"] is None:
                    c["expect"] = x
    return {"networks": nets, "checks": checks}


def main(ref):
    ref = Path(ref)
    sft = (ref / "test/SegmentForcesTest.jl").read_text().splitlines()
    fem = (ref / "test/DislocationFEMTest.jl").read_text().splitlines()
    mob = (ref / "test/MobilityTest.jl").read_text().splitlines()
    out = {}

    def put(name, src, lines, a, b, scale=1.0, flatten=False):
        m = matrix(lines, a, b)
        if scale != 1.0:
            m = [[v * scale for v in r] for r in m]
        if flatten:
            m = [v for r in m for v in r]
        out[name] = {"provenance": f"{src}:{a}-{b}", "data": m}

    S = "test/SegmentForcesTest.jl"
    # PK network and its uHat (sparse: 1-based indices, values x 1000 at use)
    out["pk_uHat_index"] = {"provenance": f"{S}:66", "data": index_list(sft[65])}
    put("pk_uHat_value", S, sft, 67, 78, flatten=True)
    put("pk_coord", S, sft, 127, 129)
    put("pk_fPKTest", S, sft, 133, 164)
    # sigma_hat
    out["sigmahat_uHat_index"] = {"provenance": f"{S}:618", "data": index_list(sft[617])}
    put("sigmahat_uHat_value", S, sft, 619, 630, flatten=True)
    put("sigmahat_1", S, sft, 638, 640)
    put("sigmahat_2", S, sft, 646, 648)
    # self force (pentagons, second loop +10)
    put("self_f1", S, sft, 245, 254)
    put("self_f2", S, sft, 257, 266)
    # seg-seg: coincident pentagons, offset pentagons (+20), hexagons
    put("segseg_coincident_f1", S, sft, 298, 307)
    put("segseg_coincident_f2", S, sft, 310, 319)
    put("segseg_offset20_f1", S, sft, 344, 353)
    put("segseg_offset20_f2", S, sft, 356, 365)
    put("segseg_hexagons_f1", S, sft, 413, 424)
    put("segseg_hexagons_f2", S, sft, 427, 438)
    # field-point stress, edge and screw single segment (columns 5/6 swapped
    # in the reference's test matrices)
    F = "test/DislocationFEMTest.jl"
    put("sigmatilde_edge", F, fem, 61, 69)
    put("sigmatilde_screw", F, fem, 138, 146, scale=1.0e-03)
    # dislocation displacements at the cantilever's u-boundary nodes
    # (calc_uTilde, in place into the dofs uDofsDln)
    put("utilde_uGammaTest", F, fem, 323, 376, flatten=True)
    put("utilde_uxTest", F, fem, 387, 440, flatten=True)
    put("utilde_uyTest", F, fem, 443, 496, flatten=True)
    put("utilde_uzTest", F, fem, 499, 552, flatten=True)
    put("utilde_uTildeTest", F, fem, 564, 1139, flatten=True)
    # collision detection: minimumDistance cases and detectCollision checks
    misc = (ref / "test/MiscTest.jl").read_text().splitlines()
    top = (ref / "test/TopologyTest.jl").read_text().splitlines()
    out["min_distance_cases"] = {"provenance": "test/MiscTest.jl:82-165",
                                 "data": min_distance_cases(misc, 82, 165)}
    out["collision_cases"] = {"provenance": "test/TopologyTest.jl:21-88",
                              "data": collision_cases(top, 21, 88)}
    # whole path + gather (node force), MobilityTest
    M = "test/MobilityTest.jl"
    put("mobility_testForce", M, mob, 79, 89)
    put("mobility_testVel", M, mob, 65, 75)

    for name in ("inputs/simParams/sampleDislocationParameters.json",
                 "inputs/simParams/sampleMaterialParameters.json",
                 "test/testData/sampleFEMParameters.json",
                 "inputs/simParams/sampleIntegrationParameters.json",
                 "test/testData/sampleDislocationParameters.json",
                 "test/testData/sampleIntegrationTime.json"):
        out["file:" + name] = {"provenance": name, "data": json.loads((ref / name).read_text())}

    dst = Path(__file__).with_name("reference_vectors.json")
    dst.write_text(json.dumps(out, indent=1, ensure_ascii=False) + "\n")
    print("wrote", dst, len(out), "entries")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
