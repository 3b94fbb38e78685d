"""The surrogate header reader (clb.read_surrogate_header) on headers written here in the two
forms the reference produces: the model file (surrogateModel.hh:252-500: std::vector
tables, scalar mapminmax settings, one block per `microbeId == k` branch, inputs named by its
Fin[iM][j] lines) and the trainer's generated header (surrogate_training/trainSurrogate.py:
245-395: fixed-size arrays, x_offset[] / x_gain[], logOutput). Random weights; the reader
must return them exactly, in the layer order and orientation the evaluator uses."""
import numpy as np
import pytest

import clb


def fmt(v):
    return repr(float(v))


def vec(a):
    return "{" + ", ".join(fmt(v) for v in a) + "}"


def model_file(nets, inputs):
    """A model-file style header: one branch per microbe id."""
    out = ["#include <vector>", "void defineSurrogateModel(plb::plint microbeId, const std::vector<std::vector<double>>& Fin,",
           "    std::vector<std::vector<double>>& Fout, std::vector<double>& bioR, plb::plint mask) {",
           "  // IW1_1 = {{9, 9}} in a comment is not a table", "  const std::size_t iM = microbeId;"]
    for i, (mid, (Ws, bs, xo, xg, yg)) in enumerate(nets.items()):
        out.append(f"  {'if' if i == 0 else 'else if'} (microbeId == {mid}) {{")
        for k, (W, b) in enumerate(zip(Ws, bs)):
            name = "IW1_1" if k == 0 else f"LW{k + 1}_{k}"
            if W.shape[0] == 1 and k > 0:
                out.append(f"    static const std::vector<double> {name} = {vec(W[0])};")
            else:
                rows = ",\n      ".join(vec(r) for r in W)
                out.append(f"    static const std::vector<std::vector<double>> {name} = {{\n      {rows}\n    }};")
            if len(b) == 1 and k == len(Ws) - 1:
                out.append(f"    static const double b{k + 1} = {fmt(b[0])};")
            else:
                out.append(f"    static const std::vector<double> b{k + 1} = {vec(b)};")
        for j in range(len(xo)):
            out.append(f"    static const double x_offset{j} = {fmt(xo[j])};")
            out.append(f"    static const double x_gain{j}   = {fmt(xg[j])};")
        out += ["    static const double x_ymin    = -1;", "    static const double y_ymin   = -1;",
                f"    static const double y_gain   = {fmt(yg)};", "    static const double y_offset = 0;",
                f"    std::vector<double> Xp0({len(xo)});"]
        for j, sub in enumerate(inputs):
            out.append(f"    Xp0[{j}] = (Fin[iM][{sub}] - x_offset{j}) * x_gain{j} + x_ymin;")
        out.append("  }")
    out.append("}")
    return "\n".join(out) + "\n"


def generated(Ws, bs, xo, xg, yg, log_output):
    out = ["namespace surrogate_test {", f"static const std::size_t nInputs = {len(xo)};",
           f"static const double x_offset[{len(xo)}] = {vec(xo)};", f"static const double x_gain  [{len(xg)}] = {vec(xg)};",
           "static const double x_ymin      = -1;", "static const double y_offset = 0.125;",
           f"static const double y_gain   = {fmt(yg)};", "static const double y_ymin   = -1;",
           f"static const bool logOutput = {'true' if log_output else 'false'};"]
    for k, (W, b) in enumerate(zip(Ws, bs)):
        name = "IW1_1" if k == 0 else f"LW{k + 1}_{k}"
        out.append(f"/* layer {k + 1}: {W.shape[0]} x {W.shape[1]} */")
        out.append(f"static const double {name}[{W.shape[0]}][{W.shape[1]}] = {{")
        out += [f"    {vec(r)}," for r in W]
        out.append("};")
        out.append(f"static const double b{k + 1}[{len(b)}] = {vec(b)};")
    out.append("}  // namespace surrogate_test")
    return "\n".join(out) + "\n"


def random_net(rng, sizes):
    Ws = [rng.normal(size=(sizes[k + 1], sizes[k])) for k in range(len(sizes) - 1)]
    bs = [rng.normal(size=sizes[k + 1]) for k in range(len(sizes) - 1)]
    return Ws, bs


def flat(Ws, bs):
    return np.concatenate([np.concatenate([W.ravel(), b]) for W, b in zip(Ws, bs)])


def test_model_file_two_microbes(tmp_path):
    rng = np.random.default_rng(1)
    n0 = random_net(rng, [2, 5, 4, 1]) + ([1e-3, 2e-5], [0.2, 4.0], 37.5)
    n3 = random_net(rng, [2, 3, 1]) + ([0.5, 0.25], [1.5, 2.5], 12.0)
    p = tmp_path / "surrogateModel.hh"
    p.write_text(model_file({0: n0, 3: n3}, inputs=[2, 0]))
    nets = clb.read_surrogate_header(str(p))
    assert sorted(nets) == [0, 3]
    for mid, (Ws, bs, xo, xg, yg) in ((0, n0), (3, n3)):
        net = nets[mid]
        assert net["widths"] == [W.shape[0] for W in Ws] and net["n_in"] == 2
        assert net["in_sub"] == [2, 0]
        np.testing.assert_array_equal(net["params"], flat(Ws, bs))
        assert net["x_offset"] == xo and net["x_gain"] == xg and net["y_gain"] == yg
        assert (net["x_ymin"], net["y_ymin"], net["y_offset"], net["log_output"]) == (-1.0, -1.0, 0.0, False)


@pytest.mark.parametrize("log_output", [False, True])
def test_generated_header(tmp_path, log_output):
    rng = np.random.default_rng(2)
    Ws, bs = random_net(rng, [3, 6, 6, 1])
    p = tmp_path / "surrogate_weights_test.hh"
    p.write_text(generated(Ws, bs, [0.1, 0.2, 0.3], [1.0, 2.0, 3.0], 9.5, log_output))
    nets = clb.read_surrogate_header(str(p))
    net = nets[0]
    assert net["widths"] == [6, 6, 1] and net["in_sub"] == [0, 1, 2]
    np.testing.assert_array_equal(net["params"], flat(Ws, bs))
    assert net["y_offset"] == 0.125 and net["log_output"] == log_output


def test_reader_rejects_broken_shapes(tmp_path):
    rng = np.random.default_rng(3)
    Ws, bs = random_net(rng, [2, 4, 1])
    bs[0] = bs[0][:3]  # a bias vector one short
    p = tmp_path / "bad.hh"
    p.write_text(generated(Ws, bs, [0.0, 0.0], [1.0, 1.0], 1.0, False))
    with pytest.raises(clb.ClbError):
        clb.read_surrogate_header(str(p))
