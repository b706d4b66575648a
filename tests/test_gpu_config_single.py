"""The single-VDF configuration's workload end to end on the GPU: 27^3 nodes
(3 Gauss nodes x 9 cells per dimension on [-4,4]^3), the seeded bimodal
Maxwellian plus a 5% perturbation on 10% of the cells, RK3 with dt = 0.005 for
500 steps (hipGraph replays), plain and with the FP64 conservation projection.
(Q of this VDF against the oracle: test_gpu_parity_large.)

  * plain (the reference's scheme): the density moves at the rate the
    evaluator's own Q gives, sum w Q (canonical records shifted and dropped at
    the grid's edge, collision_mod.f90:195-211, do not conserve mass), and the
    run is reported with the step where it stops being finite;
  * projected: density, momentum and energy stay within 1e-12 of their
    initial values over the 500 steps."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_single_vdf_configuration_rk3():
    import dgv0d3v_direct_collision_amd as dgv
    from dgv0d3v_direct_collision_amd.workloads import perturbed_bimodal_vdf
    M, L = 9, 4.0
    print("\n[config single] 27^3 A-array", flush=True)
    d = dgv.DGV(M, 3, (-L, L))
    try:
        d.generate_aarray(min_sens=0.0, nquad=24)
        u, v, w, gw = d.nodes_u, d.nodes_v, d.nodes_w, d.nodes_gwts
        f0 = perturbed_bimodal_vdf(u, v, w, gw, M, np.random.default_rng(20261203))
        q0 = d.EvalCollisionPeriodicAPlus_DGV(f0)
        dt = 0.005
        d_f = d.device_alloc(8 * d.nnodes)
        fr = np.zeros(d.nnodes)
        psi = np.stack([np.ones(d.nnodes), u, v, w, u * u + v * v + w * w])
        m0 = (f0 * gw) @ psi.T
        # plain: the first step moves the density by dt sum w Q to first order
        d.set_projection(False)
        d.htod(d_f, f0)
        d.rk_run_dev(d_f, dt, 1, rkmts=3, use_graph=True)
        d.synchronize()
        d.dtoh(fr, d_f)
        dn = float(np.sum(gw * fr) - m0[0])
        rate = float(np.sum(gw * q0))
        print(f"[config single] plain: density change in one step {dn:.4e}, dt sum w Q(f0) {dt * rate:.4e}",
              flush=True)
        assert rate > 0 and abs(dn - dt * rate) <= 0.05 * abs(dt * rate)
        done, diverged = 1, None
        while done < 500:
            d.rk_run_dev(d_f, dt, 50, rkmts=3, use_graph=True)
            d.synchronize()
            done += 50
            d.dtoh(fr, d_f)
            if not np.all(np.isfinite(fr)):
                diverged = done
                break
        print(f"[config single] plain RK3: {done} steps run, " +
              (f"not finite by step {diverged}" if diverged else f"density {np.sum(gw * fr):.6f}"), flush=True)
        # projected: the moments hold over all 500 steps
        d.set_projection(True)
        d.htod(d_f, f0)
        d.rk_run_dev(d_f, dt, 500, rkmts=3, use_graph=True)
        d.synchronize()
        d.dtoh(fr, d_f)
        m1 = (fr * gw) @ psi.T
        drift = np.abs(m1 - m0) / np.abs(m0).clip(min=1e-300)
        print(f"[config single] projected RK3 x 500: relative moment drift {drift}", flush=True)
        assert np.all(np.isfinite(fr))
        assert abs(m1[0] - m0[0]) <= 1e-12 * abs(m0[0])
        assert abs(m1[4] - m0[4]) <= 1e-12 * abs(m0[4])
        assert np.all(np.abs(m1[1:4] - m0[1:4]) <= 1e-12 * abs(m0[0]))
        d.set_projection(False)
        d.device_free(d_f)
    finally:
        d.close()
