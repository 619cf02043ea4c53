"""Graph converters (no GPU): GFA and the string-labelled CSR text in, the
two CSR texts of src/handlegraph_to_csr.cpp and the .adj inputs of
src/csr_to_adj.cpp out.  The expected files of the small graph below are
worked out by hand from the reference's rules: handles ordered by id, each
followed by its flip (handlegraph_to_csr.cpp:49-60), out-lists sorted
(:103-110), and the char expansion (:165-247)."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from chsga_amd import Dataset, ReadSet
from chsga_amd import native as N
from oracle import oracle as O

ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "chsga_amd" / "bin"

# segments out of id order in the file; links 1->2, 1->3 (overlap *), 2->3
GFA = "H\tVN:Z:1.0\nS\t3\tTTG\nS\t1\tACG\nS\t2\tG\nL\t1\t+\t2\t+\t0M\nL\t1\t+\t3\t+\t*\nL\t2\t+\t3\t+\t0M\nP\tp\t1+,2+,3+\t*\n"
# handles 0 = 1+ ACG, 1 = 1- CGT, 2 = 2+ G, 3 = 2- C, 4 = 3+ TTG, 5 = 3- CAA;
# 1+ -> 2+ gives 2- -> 1-, 1+ -> 3+ gives 3- -> 1-, 2+ -> 3+ gives 3- -> 2-
STRING_CSR = "6\n6\nACG CGT G C TTG CAA \n0 2 2 3 4 4 6 \n2 4 4 1 1 3 \n"
# chars: 0-2 ACG, 3-5 CGT, 6 G, 7 C, 8-10 TTG, 11-13 CAA; the last char of a
# string takes its string's out-edges to the first chars of the targets
CHAR_CSR = ("14\n14\nA C G C G T G C T T G C A A \n0 1 2 4 5 6 6 7 8 9 10 10 11 12 14 \n"
            "1 2 6 8 4 5 8 3 9 10 12 13 3 7 \n")


def _files(tmp_path, text=GFA, name="g"):
    p = tmp_path / f"{name}.gfa"
    p.write_text(text)
    return p


def test_gfa_lists_both_orientations_like_a_handle_graph(tmp_path):
    ds = Dataset.load_gfa(str(_files(tmp_path)))
    ds.write_csr(str(tmp_path / "s.csr"), str(tmp_path / "c.csr"))
    assert (tmp_path / "s.csr").read_text() == STRING_CSR
    assert (tmp_path / "c.csr").read_text() == CHAR_CSR
    a = ds.arrays()
    assert ds.labels() == b"ACGCGTGCTTGCAA"
    assert a["str_off"].tolist() == [0, 3, 6, 7, 8, 11, 14]


def test_inversion_link_gives_the_flipped_edge(tmp_path):
    ds = Dataset.load_gfa(str(_files(tmp_path, "S\t1\tAC\nS\t2\tGG\nL\t1\t+\t2\t-\t0M\n")))
    a = ds.arrays()
    # 1+ = 0 (AC), 1- = 1 (GT), 2+ = 2 (GG), 2- = 3 (CC): 1+ -> 2- and 2+ -> 1-
    assert ds.labels() == b"ACGTGGCC"
    out = {s: a["str_out_dst"][a["str_out_off"][s]:a["str_out_off"][s + 1]].tolist() for s in range(4)}
    assert out == {0: [3], 1: [], 2: [1], 3: []}


def test_drivers_chain_to_the_consga_inputs(tmp_path):
    """gfa_to_csr, then csr_to_adj on both CSR texts (the reference's tool
    chain), loads as the same dataset as the GFA and the string CSR."""
    g = _files(tmp_path)
    s_csr, c_csr = tmp_path / "g.str.csr", tmp_path / "g.char.csr"
    subprocess.run([str(BIN / "gfa_to_csr"), str(g), str(s_csr), str(c_csr)], check=True)
    assert s_csr.read_text() == STRING_CSR and c_csr.read_text() == CHAR_CSR
    pre = str(tmp_path / "g")
    subprocess.run([str(BIN / "csr_to_adj"), str(s_csr), pre + ".str.adj", pre + ".labels", pre + ".label_offsets"],
                   check=True)
    subprocess.run([str(BIN / "csr_to_adj"), str(c_csr), pre + ".char.adj", pre + ".char.labels",
                    pre + ".char.label_offsets"], check=True)
    assert open(pre + ".label_offsets").read() == "0 3 6 7 8 11 14\n"
    assert open(pre + ".labels").read() == "ACGCGTGCTTGCAA" == open(pre + ".char.labels").read()
    lines = open(pre + ".char.adj").read().split("\n")
    assert lines[:3] == ["AdjacencyGraph", "14", "14"] and lines[-2] == "1.0"
    a = Dataset.load(pre + ".char.adj", pre + ".str.adj", None, pre + ".labels", pre + ".label_offsets").arrays()
    b = Dataset.load_gfa(str(g)).arrays()
    c = Dataset.load_string_csr(str(s_csr)).arrays()
    for k in ("labels", "str_off", "str_out_off", "str_out_dst", "in_off", "in_src"):
        assert np.array_equal(a[k], b[k]) and np.array_equal(a[k], c[k]), k


def test_both_strands_align_with_score_zero(tmp_path):
    """A read spelling the forward path 1+ 2+ 3+ and its reverse complement
    (the path 3- 2- 1-) both align exactly; a read on no path does not."""
    ds = Dataset.load_gfa(str(_files(tmp_path)))
    reads = ReadSet.from_list([b"ACGGTTG", b"CAACCGT", b"ACGTTG", b"TTTTTTT"])
    sc = O.consga_scores(ds.desc(), reads)
    assert sc[0] == 0 and sc[1] == 0 and sc[2] == 0 and sc[3] > 0


def test_generated_graph_survives_gfa(tmp_path):
    """A generated string graph written as GFA (segment i + 1 = string i)
    comes back with its forward strands as the even string vertices."""
    ds = Dataset.generate(seed=12, backbone_bp=3000, snp_rate=1 / 25, indel_rate=1 / 150, indel_len=(1, 12))
    a = ds.arrays()
    lab = ds.labels()
    n = a["n_s"]
    lines = [f"S\t{i + 1}\t{lab[a['str_off'][i]:a['str_off'][i + 1]].decode()}" for i in range(n)]
    for i in range(n):
        for j in a["str_out_dst"][a["str_out_off"][i]:a["str_out_off"][i + 1]]:
            lines.append(f"L\t{i + 1}\t+\t{int(j) + 1}\t+\t0M")
    # file order shuffled: ids, not positions, fix the string order
    rng = np.random.default_rng(3)
    lines = [lines[k] for k in rng.permutation(len(lines))]
    g = _files(tmp_path, "\n".join(lines) + "\n", "gen")
    ds2 = Dataset.load_gfa(str(g))
    b = ds2.arrays()
    lab2 = ds2.labels()
    assert b["n_s"] == 2 * n
    for i in range(n):
        assert lab2[b["str_off"][2 * i]:b["str_off"][2 * i + 1]] == lab[a["str_off"][i]:a["str_off"][i + 1]]
        out = b["str_out_dst"][b["str_out_off"][2 * i]:b["str_out_off"][2 * i + 1]].tolist()
        assert out == [2 * int(j) for j in a["str_out_dst"][a["str_out_off"][i]:a["str_out_off"][i + 1]]]
    reads = ReadSet.from_list([lab[100:160], lab[1000:1090]])
    assert (O.consga_scores(ds2.desc(), reads) == O.consga_scores(ds.desc(), reads)).all()


@pytest.mark.parametrize("text,msg", [
    ("S\t1\tAC\nS\t2\tGG\nL\t1\t+\t2\t+\t3M\n", "blunt"),
    ("S\t1\tAC\nL\t1\t+\t9\t+\t0M\n", "unknown segment"),
    ("S\t1\t*\n", "no sequence"),
    ("S\t1\tAC\nS\t1\tGG\n", "duplicate"),
    ("S\t1\tAC\nL\t1\tx\t1\t+\n", "orientation"),
])
def test_bad_gfa_is_rejected(tmp_path, text, msg):
    with pytest.raises(N.NativeError, match=msg):
        Dataset.load_gfa(str(_files(tmp_path, text)))


def test_bad_string_csr_is_rejected(tmp_path):
    p = tmp_path / "bad.csr"
    p.write_text("2\n1\nAC G\n0 1 2\n1\n")  # last offset 2 != m = 1
    with pytest.raises(N.NativeError, match="offsets"):
        Dataset.load_string_csr(str(p))
    p.write_text("2\n1\nAC G\n0 1 1\n")
    with pytest.raises(N.NativeError, match="truncated"):
        Dataset.load_string_csr(str(p))
