"""The MATLAB-format benchmark driver on the GPU: output lines parse back
into (P, Q, R, steps, ms) points with the enum names of the reference."""
import pytest

pytestmark = pytest.mark.gpu


def test_benchmark_set_round_trip(fm, capsys):
    from fastmm import benchmark as bm
    lines = bm.benchmark_set("FAST424_26_206", [(512, 256, 512), (600, 300, 700)], [0, 1],
                             trials=3)
    data = bm.parse("\n".join(lines))
    assert set(data) == {"FAST424_26_206_0", "FAST424_26_206_1"}
    for name, pts in data.items():
        assert [p[:4] for p in pts] == [(512, 256, 512, int(name[-1])),
                                        (600, 300, 700, int(name[-1]))]
        assert all(p[4] > 0 for p in pts)
    bm.main(["--alg", "strassen", "--steps", "1", "--sizes", "256,128x64x96", "--trials", "1"])
    out = capsys.readouterr().out
    pts = bm.parse(out)["STRASSEN_1"]
    assert [p[:4] for p in pts] == [(256, 256, 256, 1), (128, 64, 96, 1)]


def test_benchmark_leaf_modes(fm, capsys):
    from fastmm import benchmark as bm
    bm.main(["--alg", "strassen", "--steps", "2", "--sizes", "512", "--trials", "1", "--fp32",
             "--mode", "split_bf16"])
    bm.main(["--alg", "smirnov333_23_139", "--steps", "1", "--sizes", "576", "--trials", "1",
             "--mode", "ozaki_int8"])
    data = bm.parse(capsys.readouterr().out)
    assert [p[:4] for p in data["STRASSEN_2"]] == [(512, 512, 512, 2)]
    assert data["STRASSEN_2"][0][4] > 0
    (name,) = [k for k in data if k != "STRASSEN_2"]
    assert data[name][0][:4] == (576, 576, 576, 1) and data[name][0][4] > 0
