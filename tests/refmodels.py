"""The reference's own test and example models, rebuilt through a builder.

``builder`` is any object with the Rustb builder surface (``tb_model``-style
constructor ``make(lat, orb, spin)`` returning a model with ``set_hop``,
``add_hop``, ``set_onsite``, ``add_onsite``); the oracle's ``Model`` and the
product's ``rustb_amd.Model`` both qualify, so the same recipe feeds both.
"""
import math

import numpy as np

SQ3 = math.sqrt(3.0)
HEX_LAT = [[1.0, 0.0], [0.5, SQ3 / 2.0]]
HEX_ORB = [[1.0 / 3.0, 1.0 / 3.0], [2.0 / 3.0, 2.0 / 3.0]]


def gen_v_model(make, t=1.0, delta=0.0):
    """test_gen_v model (lib.rs:671-700)."""
    m = make(HEX_LAT, HEX_ORB, False)
    m.set_onsite([-delta, delta])
    for R in ([0, 0], [-1, 0], [0, -1]):
        m.set_hop(t, 0, 1, R)
    return m


def haldane(make, t2_imag=-1.0, t=-1.0, delta=0.7):
    """Haldane model of conductivity_test / Haldan_model / build_haldane_2d
    (lib.rs:757-785, 902-925, 2356-2376): t2 hoppings are ``t2_imag * i``."""
    m = make(HEX_LAT, HEX_ORB, False)
    m.set_onsite([-delta, delta])
    for R in ([0, 0], [-1, 0], [0, -1]):
        m.add_hop(t, 0, 1, R)
    for R in ([1, 0], [-1, 1], [0, -1]):
        m.add_hop(complex(0.0, t2_imag), 0, 0, R)
    for R in ([-1, 0], [1, -1], [0, 1]):
        m.add_hop(complex(0.0, t2_imag), 1, 1, R)
    return m


def graphene_soc_example(make):
    """Spinful graphene of examples/graphene/main.rs:11-32."""
    t1, soc, delta, J = -2.85, 0.2, 0.5, 0.1
    lat = [[SQ3, -1.0], [SQ3, 1.0]]
    orb = [[0.0, 0.0], [1.0 / 3.0, 1.0 / 3.0]]
    m = make(lat, orb, True)
    m.add_onsite([delta, -delta], "Z")
    m.add_onsite([J, J], "Z")
    m.add_hop(t1, 0, 1, [0, 0])
    m.add_hop(t1, 0, 1, [-1, 0])
    m.add_hop(t1, 0, 1, [0, -1])
    for R in ([1, 0], [0, 1], [1, -1]):
        m.add_hop(1j * soc, 0, 0, R, "Z")
        m.add_hop(-1j * soc, 1, 1, R, "Z")
    return m


def graphene_config1(make, t=-2.85):
    """Spinless graphene of the DOS workload: lattice rows (sqrt3,-1),(sqrt3,1),
    orbitals (0,0),(1/3,1/3), NN hopping t at R=(0,0),(-1,0),(0,-1)."""
    lat = [[SQ3, -1.0], [SQ3, 1.0]]
    orb = [[0.0, 0.0], [1.0 / 3.0, 1.0 / 3.0]]
    m = make(lat, orb, False)
    for R in ([0, 0], [-1, 0], [0, -1]):
        m.add_hop(t, 0, 1, R)
    return m


def qwz_stacked_3d(make, m0=-1.0):
    """Stacked 2D QWZ of build_qwz_stacked_3d (lib.rs:3195-3217):
    H = sin kx sx + sin ky sy + (m + cos kx + cos ky) sz, kz-independent."""
    m = make(np.eye(3), [[0.0, 0.0, 0.0], [0.0, 0.0, 0.0]], False)
    m.add_onsite([m0, -m0])
    m.add_hop(complex(0.0, -0.5), 0, 1, [1, 0, 0])
    m.add_hop(complex(0.0, 0.5), 0, 1, [-1, 0, 0])
    m.add_hop(-0.5, 0, 1, [0, 1, 0])
    m.add_hop(0.5, 0, 1, [0, -1, 0])
    for R in ([1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0]):
        m.add_hop(0.5, 0, 0, R)
        m.add_hop(-0.5, 1, 1, R)
    return m
