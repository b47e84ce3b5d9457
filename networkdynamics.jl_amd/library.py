"""Component library: the reference's test/ComponentLibrary.jl and
benchmark/benchmark_models.jl constructors, plus two build-defined kinds
(first-order load vertex, static line with phase shift).

Every model here is one of the equations the fused HIP kernels implement.
"""
from . import components as c


# test/ComponentLibrary.jl:11-13
def diffusion_edge():
    return c.EdgeModel(g=c.AntiSymmetric(c.diffusionedge_weighted), outdim=1, psym=[("p", 1.0)],
                       name="diff_edge")


# test/ComponentLibrary.jl:26-28
def diffusion_edge_fid():
    return c.EdgeModel(g=c.diffusionedge_fid, outdim=1, psym=[("p", 1.0)], name="diff_edge_fid")


# test/ComponentLibrary.jl:35-40
def diffusion_odeedge():
    return c.EdgeModel(f=c.diffusion_dedge, sym=["e_dst", "e_src"], psym=[("tau", 100.0)],
                       g=c.Fiducial(dst=1, src=2), name="diff_edge_ode")


# One-state dynamic line: de = 1/tau * (sin(v_s - v_d) - e), the dst half of
# test/ComponentLibrary.jl:30-34, its output a wrapped StateMask (component_functions.jl:117-175):
# wrapper "anti" (y_src = -e, the default), "sym" (y_src = e) or "directed" (no src output).
def ode_lag_edge(wrapper="anti", tau=1.0, name="lag_line"):
    w = dict(anti=c.AntiSymmetric, sym=c.Symmetric, directed=c.Directed)[wrapper]
    return c.EdgeModel(f=c.ode_lag_f, sym=["e"], psym=[("tau", tau)], g=w(1), name=name)


# test/external_inputs_test.jl:45-53: de[1] = ext[1], output AntiSymmetric(1)
def ext_driven_edge(extin, wrapper="anti", name="e_noff"):
    w = dict(anti=c.AntiSymmetric, sym=c.Symmetric, directed=c.Directed)[wrapper]
    return c.EdgeModel(f=c.feext, sym=["a"], g=w(1), extin=extin, name=name)


# docs/examples/getting_started_with_network_dynamics.jl:218-221: d independent diffusions
# on one graph -- vertices with d states and outputs 1:d (dv .= esum), edges whose d
# outputs are v_s .- v_d (AntiSymmetric)
def diffusion_vertex_multi(d, sym=None, name="diffusion_vertex_d"):
    return c.VertexModel(f=c.diffusionvertex_vec, dim=d, sym=sym, g=c.StateMask(range(1, d + 1)), name=name)


def diffusion_edge_multi(d, outsym=None, name="diffusion_edge_d"):
    return c.EdgeModel(g=c.AntiSymmetric(c.diffusionedge), outdim=d, outsym=outsym, name=name)


# test/ComponentLibrary.jl:46
def diffusion_vertex():
    return c.VertexModel(f=c.diffusionvertex, dim=1, g=1, name="diffusion_vertex")


# test/ComponentLibrary.jl:54-57
def kuramoto_edge(name="kuramoto_edge"):
    return c.EdgeModel(g=c.AntiSymmetric(c.kuramoto_edge_g), outsym=["P"], psym=["K"], name=name)


# test/ComponentLibrary.jl:64-67
def kuramoto_second(name="kuramoto_second"):
    return c.VertexModel(f=c.kuramoto_inertia, sym=[("delta", 0.0), ("omega", 0.0)],
                         psym=[("M", 1.0), ("D", 0.1), ("Pm", 1.0)], g=c.StateMask(1), name=name)


# test/ComponentLibrary.jl:72
def kuramoto_first():
    return c.VertexModel(f=c.kuramoto_vertex, sym=["theta"], psym=["omega"], g=1, name="kuramoto_first")


# build-defined: tau * dtheta = P + sum of edge inputs; p = [tau, P].  With
# mass_matrix=0 the state is algebraic: 0 = (P + sum of edge inputs) / tau.
def load_vertex(name="load", mass_matrix=None):
    return c.VertexModel(f=c.load_vertex_f, sym=["theta"], psym=["tau", "P"], g=1, name=name,
                         mass_matrix=mass_matrix)


# build-defined: y_dst = B * sin(theta_src - theta_dst - phi), AntiSymmetric; p = [B, phi]
def static_line_edge(name="static_line"):
    return c.EdgeModel(g=c.AntiSymmetric(c.static_line_g), outsym=["P"], psym=["B", ("phi", 0.0)], name=name)


# docs/src/external_inputs.md:62-99: the resistor, the capacitor load with a periodic
# current sink, and the PI-controlled current source that reads the load's voltage
# through an external input
def resistor(name="resistor"):
    return c.EdgeModel(g=c.AntiSymmetric(c.resistor_g), outsym=["i"], psym=[("R", 0.1)], name=name)


def capacitor_load(name="load"):
    return c.VertexModel(f=c.f_load, g=1, sym=[("V", 0.5)], psym=[("C", 1.0)], name=name)


def pi_source(extin, name="source"):
    return c.VertexModel(f=c.f_source, g=1, sym=[("V", 0.5), ("xi", 0.0)],
                         psym=[("C", 1.0), ("Vref", 1.0), ("Ki", 0.5), ("Kp", 10.0)], extin=extin, name=name)


# test/loopback_test.jl:53-131: a capacitor hub with resistor and inductor injectors on
# LoopbackConnections, the voltage source feeding it through a resistor, and the
# single-vertex CRL model the injector network must reproduce
def capacitor(name="hub"):
    return c.VertexModel(f=c.capacitor_f, g=1, sym=[("v", 0.0)], psym=[("C", 1.0)], name=name)


def voltage_source(name="vs"):
    return c.VertexModel(g=c.voltage_source_g, psym=[("V", 1.0)], name=name)


def resistor_injector(name="R_injector"):
    return c.VertexModel(g=c.resistor_injector_g, psym=[("R", 100.0)], name=name)


def inductor_injector(name="L_injector"):
    return c.VertexModel(f=c.inductor_injector_f, g=1, sym=[("i", 0.0)], psym=[("L", 0.1)], name=name)


def crl(name="CRL_vertex"):
    return c.VertexModel(f=c.crl_f, g=1, sym=[("v", 0.0), ("iL", 0.0)],
                         psym=[("C", 1.0), ("R", 100.0), ("L", 0.1)], name=name)


def loopback_connection(name="loopback"):
    """LoopbackConnection (src/post_utils.jl:110-190): a Directed edge from an injector to
    its hub; the hub receives -1 times the injector's output and the injector's input is
    the hub's output."""
    return c.EdgeModel(g=c.Directed(c.loopback_g), outdim=1, name=name)


# docs/examples/directed_and_weighted_graphs.jl:89-101
def fhn_vertex(a=0.5, eps=0.05, name="fhn"):
    return c.VertexModel(f=c.fhn_electrical_vertex, g=1, sym=["u", "v"], psym=[("a", a), ("eps", eps)], name=name)


def electrical_edge(sigma=0.5, name="electrical_edge"):
    return c.EdgeModel(g=c.Directed(c.electrical_edge_g), outdim=1, psym=["weight", ("sigma", sigma)], name=name)


# docs/examples/stress_on_truss.jl:45-88: free nodes (velocity and position; the beam forces
# are their inputs), fixed nodes (the position is the parameter) and beams
def truss_free_vertex(M=10.0, gamma=200.0, g=9.81, name="free"):
    return c.VertexModel(f=c.truss_free_f, g=c.StateMask([3, 4]), sym=[("vx", 0.0), ("vy", 0.0), "x", "y"],
                         psym=[("M", M), ("gamma", gamma), ("g", g)], insym=["Fx", "Fy"], name=name)


def truss_fixed_vertex(name="fixed"):
    return c.VertexModel(g=c.truss_fixed_g, psym=["xfix", "yfix"], name=name)


def beam(K=0.5e6, name="beam"):
    return c.EdgeModel(g=c.AntiSymmetric(c.beam_g), psym=[("K", K), "L"], outsym=["Fx", "Fy"], name=name)


# docs/examples/heterogeneous_system.jl:75-105: a vertex whose output is its parameter
# (NoFeedForward g(out, p) = p, the voltage source's equation) and the Kuramoto vertex with
# inertia and unit damping (benchmark/benchmark_models.jl:37-41 with psym omega0)
def static_vertex(theta_fix=0.0, name="static"):
    return c.VertexModel(g=c.voltage_source_g, psym=[("theta_fix", theta_fix)], name=name)


def kuramoto_inertia(name="inertia"):
    return c.VertexModel(f=c.kuramoto_inertia_unit_damping, sym=["theta", "omega"], psym=["omega0"], g=1,
                         name=name)


class bench:
    """benchmark/benchmark_models.jl:5-42."""

    @staticmethod
    def diffusion_edge():            # :9
        return c.EdgeModel(g=c.AntiSymmetric(c.diffusionedge), outdim=1, pdim=0)

    @staticmethod
    def diffusion_dedge():           # :16
        return c.EdgeModel(f=c.diffusion_dedge_fast, dim=2, pdim=0, g=c.Fiducial(dst=1, src=2))

    @staticmethod
    def diffusion_vertex():          # :22
        return c.VertexModel(f=c.diffusionvertex, dim=1, pdim=0, g=c.StateMask(1))

    @staticmethod
    def static_kuramoto_edge():      # :30
        return c.EdgeModel(g=c.AntiSymmetric(c.kuramoto_edge_g), outdim=1, pdim=1)

    @staticmethod
    def kuramoto_vertex_1d():        # :35
        return c.VertexModel(f=c.kuramoto_vertex, pdim=1, sym=["theta"], g=c.StateMask(1))

    @staticmethod
    def kuramoto_vertex_2d():        # :42
        return c.VertexModel(f=c.kuramoto_inertia_unit_damping, dim=2, pdim=1, sym=["theta", "omega"],
                             g=c.StateMask(1))


class diffusion_test:
    """The models defined inline in test/diffusion_test.jl."""

    @staticmethod
    def odevertex():                 # :66
        return c.VertexModel(f=c.diffusionvertex, dim=1, g=1)

    @staticmethod
    def staticedge():                # :67
        return c.EdgeModel(g=c.AntiSymmetric(c.diffusionedge), outdim=1)

    @staticmethod
    def odeedge():                   # :108
        return c.EdgeModel(f=c.real_ode_edge, dim=2, g=c.Fiducial(2, 1), mass_matrix=0)
