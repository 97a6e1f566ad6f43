"""CPU FP64 oracle for the per-Newton-iteration assembly path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (``xkep-cae_amd/``) imports,
calls or links anything under ``oracle/``.  The only permitted users are
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py``, and there only as the checker (or the timed CPU baseline), never
as the thing measured or shipped.

The oracle is a NumPy FP64 restatement of the reference's algorithm
(gyp0bt/xkep-cae, pure Python/NumPy).  Each function cites the reference
file:line it follows.  The reference's per-pair Python loops are restated as
vectorised NumPy over a structure-of-arrays pair state; every per-pair update in
those loops is independent of the other pairs, so the vectorised form computes
the same values (the only difference is floating-point summation order in the
global scatter, which the reference itself leaves to ``np.add.at`` / scipy).

Modules
-------
beam      CR Timoshenko element f_int / analytical tangent / lumped mass
contact   broadphase, candidate filters, narrowphase, frame transport,
          hysteresis, Huber normal force, Coulomb return mapping, friction
          tangents (material + geometric)
coating   Kelvin-Voigt coating force (dashpot), stiffness, coating Coulomb
          friction and its stiffness
assembly  per-iteration composition R = f_int - (f_c + f_fric) - f_ext,
          K_T = K_beam + K_c - K_fric, Generalized-alpha effective terms
mlp       the learned prescreen's network (forward, loss, gradient, Adam)

The oracle has no Newton driver of its own: the Newton-count checks run the
product's driver (``xkep-cae_amd/solver.py``) over the oracle's assembly
(``tests/_oracle_backend.py``) against the reference solver's own
trajectories (``tests/golden/newton_*.npz``).

Pinning status of each function is listed in ``PINNING`` below and checked by
``tests/test_oracle_*.py``.
"""

PINNING = {
    # function: how it is pinned (closed forms / brute force / golden vectors)
    "beam.cr_beam_element_batch.f": "rigid-body invariance, zero at u=0, small-u limit K_lin*u, "
    "axial PL/EA, golden vectors from the reference",
    "beam.cr_beam_element_batch.K": "small-u limit T^T Ke T, symmetry, central FD of f (rtol 1e-4), "
    "golden vectors from the reference (exact algebra)",
    "beam.rotvec_to_rotmat/rotmat_to_rotvec": "exp(log(R)) = R, orthogonality, Rodrigues closed form",
    "beam.lumped_mass_diag": "closed form m/2, rho*Ip*L/2, m L^2/78 per node",
    "contact.broadphase_aabb": "brute-force O(n^2) inclusive AABB overlap; cell-size invariance",
    "contact.midpoint_prescreen": "golden vectors from the reference (DetectCandidates)",
    "contact.closest_point_segments": "dense (s,t) grid minimisation; parallel/perpendicular closed forms",
    "contact.build_contact_frame": "orthonormality, transport of t1 under rotation, golden vectors",
    "contact.update_geometry": "hysteresis truth table (g_on/g_off), golden vectors",
    "contact.huber_*": "closed-form Huber values/derivative, FD of f_c at fixed frame, golden vectors",
    "contact.friction_*": "return-map closed forms (stick/slip, |q|=mu p_n), D_t eigenvalues, "
    "golden vectors (sign quirks of the material block)",
    "coating.*": "the known answers of the reference's coating tests (linearity, action-reaction, "
    "force signs, dashpot rate, k_eff = k + c/dt, symmetry, PSD, Coulomb limit, history reset), node "
    "forces and return mapping in closed form, K_coat = FD of the coating force along a normal approach",
    "contact.update_geometry (coating)": "core gap, compression and activation in closed form",
    "coating.friction_stiffness_blocks": "central differences of the coating friction force at a fixed "
    "frame, in stick and in slip",
    "contact.strategy_detect / strategy_update_geometry": "golden vectors from the reference's "
    "ContactGeometryStrategy classes (shared-node filter on and off, with and without a config, coating config)",
    "contact.constraint_jacobian_ptp": "golden vectors from the reference's build_constraint_jacobian; central "
    "differences of the gaps; a hand-made pair with dropped entries and an inactive pair",
    "assembly.*": "golden vectors of R and K_T from the reference on a 7-wire state",
    "driver (tests/_oracle_backend.py + solver.py)": "Newton-iteration counts, cutbacks, divergence flags and "
    "active counts of the reference's own solver: frictionless, and Coulomb mu = 0.3 under two loads",
    "mlp.*": "BF16 rounding against torch, gradient against torch autograd and central differences, Adam "
    "against its closed form",
}
