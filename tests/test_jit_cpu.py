"""CPU checks of the FFCx source extraction (cudolfinx.jit, the reference's
python/cudolfinx/jit.py:15-114): bodies cut out of an FFCx-laid-out C file
equal the originals, the form_integrals order maps to the right bodies, and
the re-wrapped functions compute the same element tensors as the originals."""
import numpy as np
import pytest

from cudolfinx import jit
from mifem import problems
from mifem.dgforms import sipg_energy

from ffcx_layout import body_of, ffcx_c_file, integral_id
from test_codegen_cpu import _call, _compile


def _norm(s):
    return "\n".join(line.rstrip() for line in s.strip().splitlines() if line.strip())


def _forms():
    P = problems.elasticity_p2(2)
    D = problems.dg_plaplace(2)
    return {"elasticity_a": P.a, "elasticity_L": P.L, "sipg_J": D.a,
            "sipg_E": sipg_energy(D.V, D.extras["u"])}


def test_extract_bodies_and_order(tmp_path):
    forms = _forms()
    layout = {k: [(integral_id(I.name), body_of(I.source)) for I in f.integrals]
              for k, f in forms.items()}
    text = ffcx_c_file(layout)
    path = tmp_path / "ffcx_forms.c"
    path.write_text(text)
    for k, f in forms.items():
        bodies, order = jit.get_tabulate_tensor_sources(str(path), form_name=k)
        assert len(order) == len(f.integrals)
        for I, pos in zip(f.integrals, order):
            bid, body = bodies[pos]
            assert bid == integral_id(I.name)
            assert _norm(body).endswith(_norm(body_of(I.source)))
    # default: the last form in the file, as the reference; C text works as well
    b1, o1 = jit.get_tabulate_tensor_sources(text)
    last = list(forms)[-1]
    assert [b1[i][0] for i in o1] == [integral_id(I.name) for I in forms[last].integrals]
    # every definition once, prototypes skipped, comment / string braces ignored
    assert len(b1) == len({integral_id(I.name) for f in forms.values() for I in f.integrals})


def test_wrapped_functions_compute_the_same_element_tensors(tmp_path):
    forms = _forms()
    text = ffcx_c_file({k: [(integral_id(I.name), body_of(I.source)) for I in f.integrals]
                        for k, f in forms.items()})
    rng = np.random.default_rng(4)
    for k, f in forms.items():
        wrapped, order = jit.get_wrapped_tabulate_tensors(text, form_name=k)
        for I, pos in zip(f.integrals, order):
            name, src = wrapped[pos]
            assert name == I.name and src.startswith("void " + name + "(")
            new = _compile("#define restrict __restrict__\n" + src, tmp_path, f"w_{k}_{name}")
            old = _compile(I.source, tmp_path, f"o_{k}_{name}")
            ncell = 2 if I.type == "interior_facet" else 1
            mesh = f.mesh
            xc = rng.uniform(0, 1, (ncell * mesh.nv_per_cell, 3))
            xc[0:4] = [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]]
            if ncell == 2:
                xc[4:8] = [[1, 1, 1], [1, 0, 0], [0, 1, 0], [0, 0, 1]]
            w = rng.uniform(1, 2, 64)
            c = rng.uniform(1, 2, 16)
            eli = [0, 0]
            nA = 4096
            a = _call(new, name, 0, w, c, xc, eli, nA)
            b = _call(old, name, 0, w, c, xc, eli, nA)
            assert np.array_equal(a, b), (k, name)


def test_errors():
    with pytest.raises(ValueError):
        jit.get_tabulate_tensor_sources("int main(void) { return 0; }\n")
    with pytest.raises(NotImplementedError):
        jit.get_wrapped_tabulate_tensors("void tabulate_tensor_integral_a(double* A) { }\n",
                                         scalar_type="float")
    text = ffcx_c_file({"f": [("a1", "A[0] = 1.0;")]})
    with pytest.raises(KeyError):
        jit.get_tabulate_tensor_sources(text, form_name="g")
