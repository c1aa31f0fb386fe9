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
    # whole path + gather (node force), MobilityTest
    M = "test/MobilityTest.jl"
    put("mobility_testForce", M, mob, 79, 89)
    put("mobility_testVel", M, mob, 65, 75)

    for name in ("inputs/simParams/sampleDislocationParameters.json",
                 "inputs/simParams/sampleMaterialParameters.json",
                 "test/testData/sampleFEMParameters.json",
                 "inputs/simParams/sampleIntegrationParameters.json",
                 "test/testData/sampleIntegrationTime.json"):
        out["file:" + name] = {"provenance": name, "data": json.loads((ref / name).read_text())}

    dst = Path(__file__).with_name("reference_vectors.json")
    dst.write_text(json.dumps(out, indent=1, ensure_ascii=False) + "\n")
    print("wrote", dst, len(out), "entries")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
