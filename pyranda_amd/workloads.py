"""Workload definitions: Taylor-Green EOM strings and seeded synthetic fields.

Operator microbenchmark field (256^3 periodic): a sum of 64 random Fourier
modes with integer wavevectors 1 <= |k| <= 32 drawn uniformly in the ball by
rejection, amplitudes |k|^(-5/3) and uniform phases, seed 1234. Its exact
derivatives/filters are known mode by mode, which gives size-independent
checks at full size.
"""
import math

import numpy as np
import torch


def fourier_modes(n_modes=64, kmax=32, seed=1234):
    rng = np.random.default_rng(seed)
    ks = []
    while len(ks) < n_modes:
        k = rng.integers(-kmax, kmax + 1, size=3)
        r = math.sqrt(float((k * k).sum()))
        if 1.0 <= r <= kmax:
            ks.append(k)
    ks = np.array(ks, dtype=np.int64)
    amps = np.sqrt((ks * ks).sum(1).astype(np.float64)) ** (-5.0 / 3.0)
    phis = rng.uniform(0.0, 2.0 * np.pi, n_modes)
    return ks, amps, phis


def periodic_axis(n, device, dtype=torch.float64):
    """Node coordinates of xdom=(0, 2*pi*(n-1)/n, n, periodic=True)."""
    return torch.arange(n, device=device, dtype=dtype) * (2.0 * math.pi / n)


def fourier_field(n, device, modes=None, weights=None, dtype=torch.float64):
    """sum_m w_m a_m cos(k_m . x + phi_m) on an n^3 periodic grid, Fortran layout.

    `weights` (per mode) lets callers build exact operator results, e.g. the
    transfer function of an operator evaluated at each mode's wavenumber.
    Returns a (n,n,n) view with x fastest.
    """
    ks, amps, phis = modes if modes is not None else fourier_modes()
    x = periodic_axis(n, device, dtype)
    out = torch.zeros((n, n, n), device=device, dtype=dtype)  # indexed [z, y, x]
    for m in range(len(amps)):
        a = amps[m] * (1.0 if weights is None else weights[m])
        if a == 0.0:
            continue
        kx, ky, kz = (int(v) for v in ks[m])
        arg = (kx * x)[None, None, :] + (ky * x)[None, :, None] + (kz * x)[:, None, None] + phis[m]
        out += a * torch.cos(arg)
    return out.permute(2, 1, 0)


# Taylor-Green Euler+LAD equations and initial conditions, verbatim from
# examples/TaylorGreen.py:40-106 (the reference's EOM strings are its API).
TG_EOM = """
# Primary Equations of motion here
ddt(:rho:)  =  -ddx(:rho:*:u:)            - ddy(:rho:*:v:)               - ddz(:rho:*:w:)
ddt(:rhou:) =  -ddx(:rhou:*:u: - :tauxx:) - ddy(:rhou:*:v: - :tauxy:)    - ddz(:rhou:*:w: - :tauxz:)
ddt(:rhov:) =  -ddx(:rhov:*:u: - :tauxy:) - ddy(:rhov:*:v: - :tauyy:)    - ddz(:rhov:*:w: - :tauyz:)
ddt(:rhow:) =  -ddx(:rhow:*:u: - :tauxz:) - ddy(:rhow:*:v: - :tauyz:)    - ddz(:rhow:*:w: - :tauzz:)
ddt(:Et:)   =  -ddx( (:Et: - :tauxx:)*:u: - :tauxy:*:v: - :tauxz:*:w: )  - ddy( (:Et: - :tauyy:)*:v: -:tauxy:*:u: - :tauyz:*:w:) - ddz( (:Et: - :tauzz:)*:w: - :tauxz:*:u: - :tauyz:*:v: )
# Conservative filter of the EoM
:rho:       =  fbar( :rho:  )
:rhou:      =  fbar( :rhou: )
:rhov:      =  fbar( :rhov: )
:rhow:      =  fbar( :rhow: )
:Et:        =  fbar( :Et:   )
# Update the primatives and enforce the EOS
:u:         =  :rhou: / :rho:
:v:         =  :rhov: / :rho:
:w:         =  :rhow: / :rho:
:p:         =  ( :Et: - .5*:rho:*(:u:*:u: + :v:*:v: + :w:*:w:) ) * ( :gamma: - 1.0 )
# Artificial bulk viscosity
:ux:        =  ddx(:u:)
:vy:        =  ddy(:v:)
:wz:        =  ddz(:w:)
:div:       =  :ux: + :vy: + :wz:
# Remaining cross derivatives
:uy:        =  ddy(:u:)
:uz:        =  ddz(:u:)
:vx:        =  ddx(:v:)
:vz:        =  ddz(:v:)
:wy:        =  ddy(:w:)
:wx:        =  ddx(:w:)
:enst:      = sqrt( (:uy:-:vx:)**2 + (:uz: - :wx:)**2 + (:vz:-:wy:)**2 )
:tke:       = :rho:*(:u:*:u: + :v:*:v: + :w:*:w:)
:S:         = sqrt( :ux:*:ux: + :vy:*:vy: + :wz:*:wz: + .5*((:uy:+:vx:)**2 + (:uz: + :wx:)**2 + (:vz:+:wy:)**2) )
:mu:        =  gbar( abs(ring(:S:  )) ) * :rho: * 1.0e-4
:beta:      =  gbar( abs(ring(:div:)) * :rho: )  * 7.0e-3
:taudia:    =  (:beta:-2./3.*:mu:) *:div: - :p:
:tauxx:     =  2.0*:mu:*:ux:   + :taudia:
:tauyy:     =  2.0*:mu:*:vy:   + :taudia:
:tauzz:     =  2.0*:mu:*:wz:   + :taudia:
:tauxy:     = :mu:*(:uy:+:vx:)
:tauxz:     = :mu:*(:uz:+:wx:)
:tauyz:     = :mu:*(:vz:+:wz:)
:cs:  = sqrt( :p: / :rho: * :gamma: )
:dt: = dt.courant(:u:,:v:,:w:,:cs:)*1.0
:dt: = numpy.minimum(:dt:,0.2 * dt.diff(:beta:,:rho:))
:dt: = numpy.minimum(:dt:,0.2 * dt.diff(:mu:,:rho:))
"""

TG_IC = """
:gamma: = 1.4
u0 = 1.0
p0 = 100.0
rho0 = 1.0
L = 1.0
:u: =  u0*sin(meshx/L)*cos(meshy/L)*cos(meshz/L)
:v: = -u0*cos(meshx/L)*sin(meshy/L)*cos(meshz/L)
:w: = 0.0*:u:
:p:  = p0 + rho0/16.0*( ( cos(2.*meshx/L) + cos(2.*meshy/L) ) * ( cos(2.*meshz/L) + 2.0 ) - 2.0 )
:rho: = rho0 + 0.0*:u:
:rhou: = :rho:*:u:
:rhov: = :rho:*:v:
:rhow: = :rho:*:w:
:Et:  = :p: / (:gamma:-1.0) + 0.5*:rho:*(:u:*:u: + :v:*:v: + :w:*:w:)
:cs:  = sqrt( :p: / :rho: * :gamma: )
:tke: = :rho:*(:u:*:u: + :v:*:v: + :w:*:w:)
:dt: = dt.courant(:u:,:v:,:w:,:cs:)
"""


# Kelvin-Helmholtz shear layers (McNally et al.), 2-D doubly periodic, verbatim
# from examples/KH.py:56-118 (EOM and IC strings are the reference's API);
# the dictionaries are applied by pyrandaSim.EOM / setIC as in the script.
KH_EOM = """
# Primary Equations of motion here
ddt(:rho:)  =  -ddx(:rho:*:u:)            - ddy(:rho:*:v:)
ddt(:rhou:) =  -ddx(:rhou:*:u: - :tauxx:) - ddy(:rhou:*:v: - :tauxy:)
ddt(:rhov:) =  -ddx(:rhov:*:u: - :tauxy:) - ddy(:rhov:*:v: - :tauyy:)
ddt(:Et:)   =  -ddx( (:Et: - :tauxx:)*:u: - :tauxy:*:v:  - :tx:*:kappa: )  - ddy( (:Et: - :tauyy:)*:v: -:tauxy:*:u:- :ty:*:kappa: )
# Conservative filter of the EoM
:rho:       =  fbar( :rho:  )
:rhou:      =  fbar( :rhou: )
:rhov:      =  fbar( :rhov: )
:Et:        =  fbar( :Et:   )
# Update the primatives and enforce the EOS
:u:         =  :rhou: / :rho:
:v:         =  :rhov: / :rho:
:p:         =  ( :Et: - .5*:rho:*(:u:*:u: + :v:*:v:) ) * ( gamma - 1.0 )
# Artificial bulk viscosity
:ux:        =  ddx(:u:)
:vy:        =  ddy(:v:)
:div:       =  :ux: + :vy:
# Remaining cross derivatives
:uy:        =  ddy(:u:)
:vx:        =  ddx(:v:)
:enst:      = sqrt( (:uy:-:vx:)**2 )
:tke:       = :rho:*(:u:*:u: + :v:*:v: )
:S:         = sqrt( :ux:*:ux: + :vy:*:vy:  + .5*((:uy:+:vx:)**2  ) )
:mu:        =  gbar( abs(ring(:S:  )) ) * :rho: * 1.0e-4
:beta:      =  gbar( abs(ring(:div:)) * :rho: )  * 7.0e-3
:taudia:    =  (:beta:-2./3.*:mu:) *:div: - :p:
:tauxx:     =  2.0*:mu:*:ux:   + :taudia:
:tauyy:     =  2.0*:mu:*:vy:   + :taudia:
:tauxy:     = :mu:*(:uy:+:vx:)
:T:         = :p: / (:rho: * R0 )
[:tx:,:ty:,:tz:] = grad(:T:)
:kappa:     = gbar( ring(:T:)* :rho:*cv/(:T: * :dt: ) ) * 1.0e-3
:cs:  = sqrt( :p: / :rho: * gamma )
:dt: = dt.courant(:u:,:v:,:w:,:cs:)*1.0
:dt: = numpy.minimum(:dt:,0.2 * dt.diff(:beta:,:rho:))
:dt: = numpy.minimum(:dt:,0.2 * dt.diff(:mu:,:rho:))
"""

KH_IC = """
Um = (u1-u2)/2.0
rhoM = (rho1-rho2)/2.0
:u: =                     u1-Um*exp( -(meshy-.75)/L)
:u: = where( meshy < .75, u2+Um*exp( -(.75-meshy)/L) , :u: )
:u: = where( meshy < .50, u2+Um*exp( (-meshy+.25)/L) , :u: )
:u: = where( meshy < .25, u1-Um*exp(  (meshy-.25)/L) , :u: )
:v: = 0.01*sin( 4.*pi*meshx ) + Vmean
:rho: =                     rho1-rhoM*exp( -(meshy-.75)/L)
:rho: = where( meshy < .75, rho2+rhoM*exp( -(.75-meshy)/L) , :rho: )
:rho: = where( meshy < .50, rho2+rhoM*exp( (-meshy+.25)/L) , :rho: )
:rho: = where( meshy < .25, rho1-rhoM*exp(  (meshy-.25)/L) , :rho: )
:p: += p0
# Form conserved variables
:Et: = :p:/( gamma - 1.0 ) + .5*:rho:*(:u:*:u: + :v:*:v:)
:rhou: = :rho:*:u:
:rhov: = :rho:*:v:
:cs:  = sqrt( :p: / :rho: * gamma )
:dt: = dt.courant(:u:,:v:,:w:,:cs:)*.1
"""


def kh_setup(pr, nx, ny=None, vmean=0.0, precision="fp64", name="KelvinHelmholtz", device=None,
             device_scalars=False, wide_fields="auto"):
    """A KH.py simulation on an nx x ny grid (2-D, doubly periodic, the
    script's unit box with spacing 1/n), returned after setIC. The script
    uses a square grid (KH.py:52-55); ny != nx gives the uneven grids of an
    ensemble."""
    ny = nx if ny is None else ny
    gamma, R = 5. / 3., 1.0
    cp = R / (1.0 - 1.0 / gamma)
    cv = cp - R
    mesh = """
    xdom = (0.0, LX,  NX, periodic=True)
    ydom = (0.0, LY,  NY, periodic=True)
    """.replace("LX", str(1.0 * (nx - 1.0) / nx)).replace("LY", str(1.0 * (ny - 1.0) / ny)) \
       .replace("NX", str(nx)).replace("NY", str(ny))
    kw = {"silent": True, "precision": precision, "device_scalars": device_scalars,
          "wide_fields": wide_fields}
    if device is not None:
        kw["device"] = device
    ss = pr.pyrandaSim(name, mesh, **kw)
    ss.addPackage(pr.pyrandaTimestep(ss))
    ss.EOM(KH_EOM, {"gamma": gamma, "R0": R, "cv": cv})
    ss.setIC(KH_IC, {"gamma": gamma, "u1": 0.5, "u2": -0.5, "p0": 2.5, "rho1": 1.0,
                     "rho2": 2.0, "L": 0.025, "Vmean": vmean})
    return ss


def kh_run(ss, t_final=1.5, dt_max=1.0, max_steps=None):
    """The KH.py time loop (KH.py:130-151): dt grows by at most 10% a step."""
    time = 0.0
    dt = float(ss.var("dt").data)
    steps = 0
    while t_final > time:
        time = ss.rk4(time, dt)
        dt = min(float(ss.variables["dt"].data), 1.1 * dt)
        dt = min(dt_max, dt)
        dt = min(dt, (t_final - time))
        steps += 1
        if max_steps is not None and steps >= max_steps:
            break
    return time, steps


def kh_advance(ss, st, dt_max=1.0):
    """One step of the KH.py loop with time and dt as 0-d device tensors
    (a simulation built with device_scalars=True): rk4, then dt grows by at
    most 10% and stays below dt_max. The t_final clamp of the script is left
    to the caller, which runs a fixed number of steps."""
    dt = st["dt"]
    st["time"] = ss.rk4(st["time"], dt)
    nd = torch.minimum(ss.variables["dt"].data, 1.1 * dt)
    st["dt"] = torch.clamp(nd, max=dt_max)


def ensemble_sizes(n_members=64, lo=128, hi=512, seed=2024):
    """Grid sizes (nx, ny) of an ensemble: each axis log-uniform in [lo, hi],
    rounded to a multiple of 8, from a fixed seed."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_members):
        nx, ny = (int(8 * round(math.exp(rng.uniform(math.log(lo), math.log(hi))) / 8)) for _ in range(2))
        out.append((min(max(nx, lo), hi), min(max(ny, lo), hi)))
    return out


def kh_ensemble(pr, sizes, precision="fp32", device=None):
    """KH members on the given grids, each with its own seeded-free IC (the
    script's), plus their initial driver state {time, dt}. Members store every
    field in the run's precision, as a group does (wide_fields=None)."""
    sims, state = [], []
    for i, (nx, ny) in enumerate(sizes):
        ss = kh_setup(pr, nx, ny, precision=precision, name="KH-member-%d" % i, device=device,
                      device_scalars=True, wide_fields=None)
        dt = ss.var("dt").data
        dt = dt if torch.is_tensor(dt) else torch.tensor(float(dt), dtype=torch.float64,
                                                         device=ss.PyMPI.device)
        sims.append(ss)
        state.append({"time": torch.zeros((), dtype=torch.float64, device=ss.PyMPI.device),
                      "dt": dt.to(torch.float64).clone()})
    return sims, state


# Richtmyer-Meshkov, 2-D branch of examples/RM3D.py (EOM :55-131 and IC
# :139-176 verbatim; the reference's EOM and IC strings are its API): Mach ~1.5
# shock into a perturbed heavy/light interface (rho_h/rho_l = 3), bounded x
# with BC lines and sponges, periodic y.
RM_EOM = """
# Primary Equations of motion here
ddt(:rhoYh:)  =  -ddx(:rhoYh:*:u: - :Jx:)    - ddy(:rhoYh:*:v: - :Jy:)   - ddz(:rhoYh:*:w: - :Jz:)
ddt(:rhoYl:)  =  -ddx(:rhoYl:*:u: + :Jx:)    - ddy(:rhoYl:*:v: + :Jy:)   - ddz(:rhoYl:*:w: + :Jz:)
ddt(:rhou:)   =  -ddx(:rhou:*:u: - :tauxx:)  - ddy(:rhou:*:v: - :tauxy:) - ddz(:rhou:*:w: - :tauxz:)
ddt(:rhov:)   =  -ddx(:rhov:*:u: - :tauxy:)  - ddy(:rhov:*:v: - :tauyy:) - ddz(:rhov:*:w: - :tauyz:)
ddt(:rhow:)   =  -ddx(:rhow:*:u: - :tauxz:)  - ddy(:rhow:*:v: - :tauyz:) - ddz(:rhow:*:w: - :tauzz:)
ddt(:Et:)     =  -ddx( (:Et: - :tauxx:)*:u: - :tauxy:*:v: - :tauxz:*:w: ) - ddy( (:Et: - :tauyy:)*:v: -:tauxy:*:u: - :tauyz:*:w:) - ddz( (:Et: - :tauzz:)*:w: - :tauxz:*:u: - :tauyz:*:v: )
# Conservative filter of the EoM
:rhoYh:     =  fbar( :rhoYh:  )
:rhoYl:     =  fbar( :rhoYl:  )
:rhou:      =  fbar( :rhou: )
:rhov:      =  fbar( :rhov: )
:rhow:      =  fbar( :rhow: )
:Et:        =  fbar( :Et:   )
# Update the primatives and enforce the EOS
:rho:       = :rhoYh: + :rhoYl:
:Yh:        =  :rhoYh: / :rho:
:Yl:        =  :rhoYl: / :rho:
:u:         =  :rhou: / :rho:
:v:         =  :rhov: / :rho:
:w:         =  :rhow: / :rho:
:p:         =  ( :Et: - .5*:rho:*(:u:*:u: + :v:*:v: + :w:*:w:) ) * ( :gamma: - 1.0 )
# Artificial bulk viscosity / shear viscosity
:ux:        =  ddx(:u:)
:vy:        =  ddy(:v:)
:wz:        =  ddz(:w:)
:div:       =  :ux: + :vy: + :wz:
# Remaining cross derivatives
:uy:        =  ddy(:u:)
:uz:        =  ddz(:u:)
:vx:        =  ddx(:v:)
:vz:        =  ddz(:v:)
:wy:        =  ddy(:w:)
:wx:        =  ddx(:w:)
:Yx:        =  ddx(:Yh:)
:Yy:        =  ddy(:Yh:)
:Yz:        =  ddz(:Yh:)
:enst:      = sqrt( (:uy:-:vx:)**2 + (:uz: - :wx:)**2 + (:vz:-:wy:)**2 )
:tke:       = :rho:*(:u:*:u: + :v:*:v: + :w:*:w:)
:S:         = sqrt( :ux:*:ux: + :vy:*:vy: + :wz:*:wz: + .5*((:uy:+:vx:)**2 + (:uz: + :wx:)**2 + (:vz:+:wy:)**2) )
:mu:        =  gbar( abs(ring(:S:  )) ) * :rho: * 1.0e-4
:beta:      =  gbar( abs(ring(:div:)) * :rho: )  * 7.0e-3
# Artificial species diffusivities
:Dsgs:      =  ring(:Yh:) * 1.0e-4
:Ysgs:      =  1.0e2*(abs(:Yh:) - 1.0 + abs(1.0-:Yh: ) )*gridLen**2
:adiff:     =  gbar( :rho:*numpy.maximum(:Dsgs:,:Ysgs:) / :dt: )
:Jx:        =  :adiff:*:Yx:
:Jy:        =  :adiff:*:Yy:
:Jz:        =  :adiff:*:Yz:
:taudia:    =  (:beta:-2./3.*:mu:) *:div: - :p:
:tauxx:     =  2.0*:mu:*:ux:   + :taudia:
:tauyy:     =  2.0*:mu:*:vy:   + :taudia:
:tauzz:     =  2.0*:mu:*:wz:   + :taudia:
:tauxy:     = :mu:*(:uy:+:vx:)
:tauxz:     = :mu:*(:uz:+:wx:)
:tauyz:     = :mu:*(:vz:+:wz:)
:cs:        = sqrt( :p: / :rho: * :gamma: )
# Time step control routines
:dt:        = dt.courant(:u:,:v:,:w:,:cs:)*1.0
:dt:        = numpy.minimum(:dt:,0.2 * dt.diff(:beta:,:rho:))
:dt:        = numpy.minimum(:dt:,0.2 * dt.diff(:mu:,:rho:))
:dt:        = numpy.minimum(:dt:,0.2 * dt.diff(:adiff:,:rho:))
# Add some BCs in the x-direction
bc.const(['Yh'],['xn'],0.0)
bc.const(['Yh'],['x1'],1.0)
bc.const(['Yl'],['x1'],0.0)
bc.const(['Yl'],['xn'],1.0)
bc.extrap(['rho','Et'],['x1','xn'])
bc.extrap(['u'],['x1','xn'])
bc.const(['v','w'],['x1','xn'],0.0)
# Sponge BCs
:leftBC:  = 1.0 - 0.5 * (1.0 + tanh( (meshx-1.0) / .1 ) )
:rightBC: = 0.5 * (1.0 + tanh( (meshx-8.0) / .1 ) )
:BC: = numpy.maximum(:leftBC:,:rightBC:)
:u: = gbar(:u:)*:BC: + :u:*(1.0 - :BC:)
:rho: = gbar(:rho:)*:BC: + :rho:*(1.0 - :BC:)
:Et: = gbar(:Et:)*:BC: + :Et:*(1.0 - :BC:)
:rhoYh: = :rho:*:Yh:
:rhoYl: = :rho:*:Yl:
:rhou: = :rho:*:u:
:rhov: = :rho:*:v:
:rhow: = :rho:*:w:
"""

RM_IC = """
:gamma:= 5./3.
L      =  waveLength
u0     = -291.575
p0     = 100000.0
rho0_h = 3.0
rho0_l = 1.0
rhoS   = 6.37497
uS     = -61.48754
pS     = 399995.9
Amp = 3.5 + .05*(sin(meshy*L) + cos(meshz*L) ) + .25*gbar( random3D() - .5 )
:Yh: = .5 * (1.0-tanh( sqrt(pi)*( meshx - Amp) / delta ))
:Yl: = 1.0 - :Yh:
:p:  += p0
:v: *= 0.0
:w: *= 0.0
:u: = gbar( where( meshx < 3.0, uS, u0) )
:rho: = gbar( where( meshx < 3.0, rhoS, rho0_h ) )
:rho: = where( meshx < 3.2, :rho:, rho0_h * :Yh: + rho0_l * :Yl:)
:p: = gbar( where( meshx < 3.0, pS, :p: ) )
:rhoYh: = :rho:*:Yh:
:rhoYl: = :rho:*:Yl:
:rhou: = :rho:*:u:
:rhov: = :rho:*:v:
:rhow: = :rho:*:w:
:Et:  = :p: / (:gamma:-1.0) + 0.5*:rho:*(:u:*:u: + :v:*:v: + :w:*:w:)
:cs:  = sqrt( :p: / :rho: * :gamma: )
:tke: = :rho:*(:u:*:u: + :v:*:v: + :w:*:w:)
:dt: = dt.courant(:u:,:v:,:w:,:cs:)
"""

# the LAD lines the learned closure replaces (beta, then the adiff chain)
RM_LAD_LINES = (":beta:      =  gbar( abs(ring(:div:)) * :rho: )  * 7.0e-3",
                ":Dsgs:      =  ring(:Yh:) * 1.0e-4",
                ":Ysgs:      =  1.0e2*(abs(:Yh:) - 1.0 + abs(1.0-:Yh: ) )*gridLen**2",
                ":adiff:     =  gbar( :rho:*numpy.maximum(:Dsgs:,:Ysgs:) / :dt: )")
RM_CLOSURE_LINE = "[:beta:,:adiff:] = lad.closure(:div:,:vx:,:uy:,:rho:,:Yh:,:dt:)"


def rm_setup(pr, npts=128, precision="fp64", seed=1234, closure=None, name="RM_2D", device=None, nx=None,
             closure_kw=None, single_mode=False):
    """examples/RM3D.py, 2-D branch, on an nx x npts grid (nx defaults to
    int(1.4 npts), the script's choice; x spans 2.8 pi along the shock, y
    2 pi periodic across it); the IC's random perturbation draws from numpy's
    global generator seeded with `seed`. single_mode: the interface is
    perturbed by one Fourier mode (wavenumber 4 across y, amplitude 0.05,
    phase drawn from `seed`) instead of the script's mode plus filtered
    per-cell noise, whose roughness grows with resolution (at 1024 x 512
    the noise moves the interface by about two cells from one cell to the
    next, and the first step, whose dt is the initial Courant limit, is
    unstable). With `closure` (an MLPClosure) the beta and adiff lines become
    the learned closure (plain chain as fallback); `closure_kw` are keyword
    arguments of its pyrandaLearnedLAD (sensor_factor, judge_first, ...)."""
    import numpy
    from .closure import pyrandaLearnedLAD
    nx = int(npts * 1.4) if nx is None else int(nx)
    imesh = """
    xdom = (0.0, 2.8*pi , NX, periodic=False)
    ydom = (0.0, 2*pi*FF,  Npts, periodic=True)
    zdom = (0.0, 2*pi*FF,  1, periodic=True)
    """.replace("NX", str(nx)).replace("Npts", str(npts)).replace("pi", str(numpy.pi)) \
        .replace("FF", str(float(npts - 1) / npts))
    kw = {"silent": True, "precision": precision}
    if device is not None:
        kw["device"] = device
    ss = pr.pyrandaSim(name, imesh, **kw)
    ss.addPackage(pr.pyrandaTimestep(ss))
    ss.addPackage(pr.pyrandaBC(ss))
    eom = RM_EOM
    if closure is not None:
        ss.addPackage(pyrandaLearnedLAD(ss, closure, **(closure_kw or {})))
        eom = eom.replace(RM_LAD_LINES[0], RM_CLOSURE_LINE)
        for line in RM_LAD_LINES[1:]:
            eom = eom.replace(line + "\n", "")
    ss.EOM(eom)
    numpy.random.seed(seed)
    ic = RM_IC
    if single_mode:
        phase = float(numpy.random.default_rng(seed).uniform(0.0, 2.0 * numpy.pi))
        ic = ic.replace("Amp = 3.5 + .05*(sin(meshy*L) + cos(meshz*L) ) + .25*gbar( random3D() - .5 )",
                        "Amp = 3.5 + .05*sin(meshy*L + %r)" % phase)
    ss.setIC(ic, {"waveLength": 4, "delta": 2.0 * numpy.pi / npts * 4})
    return ss


def rm_run(ss, nsteps, cfl=1.0, state=None):
    """The RM3D.py loop (:196-201): dt from the EOM's dt line, growth at most
    10% per step; returns the driver state {time, dt, dtmax} to continue."""
    if state is None:
        dt = float(ss.variables["dt"].data) * cfl
        state = {"time": 0.0, "dt": dt, "dtmax": dt * .1}
    time, dt, dtmax = state["time"], state["dt"], state["dtmax"]
    for _ in range(nsteps):
        time = ss.rk4(time, dt)
        dt = float(ss.variables["dt"].data) * cfl
        dt = min(dt, dtmax * 1.1)
        dtmax = dt * 1.0
    return {"time": time, "dt": dt, "dtmax": dtmax}


def kh_group(pr, sizes, precision="fp32", device=None):
    """The KH members of `sizes` as ONE pyrandaGroupSim (shared allocation,
    one launch per operator class for all members) plus its driver state
    {time, dt}: (M,) FP64 device tensors, one value per member."""
    from .group import pyrandaGroupSim
    gamma, R = 5. / 3., 1.0
    cp = R / (1.0 - 1.0 / gamma)
    cv = cp - R
    meshes = []
    for nx, ny in sizes:
        meshes.append("""
        xdom = (0.0, LX,  NX, periodic=True)
        ydom = (0.0, LY,  NY, periodic=True)
        """.replace("LX", str(1.0 * (nx - 1.0) / nx)).replace("LY", str(1.0 * (ny - 1.0) / ny))
                      .replace("NX", str(nx)).replace("NY", str(ny)))
    kw = {"precision": precision}
    if device is not None:
        kw["device"] = device
    g = pr.pyrandaGroupSim("KH-group", meshes, **kw)
    g.addPackage(pr.pyrandaTimestep(g))
    g.EOM(KH_EOM, {"gamma": gamma, "R0": R, "cv": cv})
    g.setIC(KH_IC, {"gamma": gamma, "u1": 0.5, "u2": -0.5, "p0": 2.5, "rho1": 1.0,
                    "rho2": 2.0, "L": 0.025, "Vmean": 0.0})
    dt = g.var("dt").data.to(torch.float64).clone()
    return g, {"time": torch.zeros_like(dt), "dt": dt}


# --- Kelvin-Helmholtz shear layers between walls (ensemble workload) ---------
# One shear layer at y = 1/2 in a unit box, periodic in x with walls at y = 0
# and 1: velocity jump 1 (u = +-1/2), density ratio 2 (rho 1 below, 2 above),
# Mach 0.3 on the velocity jump against the light side's sound speed
# (p0 = rho1 (1/0.3)^2 / gamma), a seeded multimode v perturbation of
# amplitude 1e-2 localised at the interface. The EOM is KH.py's Euler + LAD
# set (KH_EOM) followed by the wall conditions of examples/euler.py:78-82
# (extrapolated rho and Et, u = v = 0 on the wall planes).
KH_WALL_BC = """
# Walls at y = 0 and y = 1
bc.extrap(['rho','Et'],['y1','yn'])
bc.const(['u','v'],['y1','yn'],0.0)
"""


def kh_wall_ic(seed=2024, modes=6, amp=1e-2, delta=0.025, mach=0.3):
    """IC string of the wall-bounded KH layer with a multimode perturbation
    drawn from `seed` (integer wavenumbers 1..8 in x, normalised so that the
    sum's largest value is `amp`)."""
    rng = np.random.default_rng(seed)
    ks = rng.choice(np.arange(1, 9), size=modes, replace=False)
    a = rng.uniform(0.5, 1.0, modes)
    ph = rng.uniform(0.0, 2.0 * math.pi, modes)
    xs = np.linspace(0.0, 1.0, 4097)
    norm = np.abs(sum(a[m] * np.sin(2 * math.pi * ks[m] * xs + ph[m]) for m in range(modes))).max()
    terms = " + ".join("%r*sin(%r*meshx + %r)" % (float(a[m] / norm), float(2 * math.pi * ks[m]), float(ph[m]))
                       for m in range(modes))
    gamma = 5. / 3.
    p0 = 1.0 * (1.0 / mach) ** 2 / gamma
    return """
:u:   = 0.5*tanh( (meshy - 0.5)/DELTA )
:rho: = 1.5 + 0.5*tanh( (meshy - 0.5)/DELTA )
:v:   = AMP*( TERMS )*exp( -((meshy - 0.5)/(4.0*DELTA))**2 )
:p:   = P0 + 0.0*meshx
:Et: = :p:/( gamma - 1.0 ) + .5*:rho:*(:u:*:u: + :v:*:v:)
:rhou: = :rho:*:u:
:rhov: = :rho:*:v:
:cs:  = sqrt( :p: / :rho: * gamma )
:dt: = dt.courant(:u:,:v:,:w:,:cs:)*.1
""".replace("DELTA", repr(delta)).replace("AMP", repr(amp)).replace("TERMS", terms).replace("P0", repr(p0)) \
        .replace("gamma", repr(gamma))


def _kh_wall_mesh(nx, ny):
    return """
    xdom = (0.0, LX,  NX, periodic=True)
    ydom = (0.0, 1.0, NY)
    """.replace("LX", str(1.0 * (nx - 1.0) / nx)).replace("NX", str(nx)).replace("NY", str(ny))


def kh_wall_setup(pr, nx, ny, precision="fp32", device=None, device_scalars=True, wide_fields=None, seed=2024):
    """One wall-bounded KH layer on an nx x ny grid (after setIC), stored like
    a group member (wide_fields=None: every field in the run's precision)."""
    gamma, R = 5. / 3., 1.0
    cp = R / (1.0 - 1.0 / gamma)
    kw = {"silent": True, "precision": precision, "device_scalars": device_scalars, "wide_fields": wide_fields}
    if device is not None:
        kw["device"] = device
    ss = pr.pyrandaSim("KH-wall-%dx%d" % (nx, ny), _kh_wall_mesh(nx, ny), **kw)
    ss.addPackage(pr.pyrandaTimestep(ss))
    ss.addPackage(pr.pyrandaBC(ss))
    ss.EOM(KH_EOM + KH_WALL_BC, {"gamma": gamma, "R0": R, "cv": cp - R})
    ss.setIC(kh_wall_ic(seed))
    return ss


def kh_wall_ensemble(pr, sizes, precision="fp32", device=None, seed=2024):
    """Wall-bounded KH members run one by one, plus their driver state."""
    sims, state = [], []
    for nx, ny in sizes:
        ss = kh_wall_setup(pr, nx, ny, precision=precision, device=device, seed=seed)
        dt = ss.var("dt").data
        dt = dt if torch.is_tensor(dt) else torch.tensor(float(dt), dtype=torch.float64, device=ss.PyMPI.device)
        sims.append(ss)
        state.append({"time": torch.zeros((), dtype=torch.float64, device=ss.PyMPI.device), "dt": dt.clone()})
    return sims, state


def kh_wall_group(pr, sizes, precision="fp32", device=None, seed=2024):
    """The wall-bounded KH members as ONE pyrandaGroupSim plus its driver
    state {time, dt} ((M,) FP64 device tensors)."""
    gamma, R = 5. / 3., 1.0
    cp = R / (1.0 - 1.0 / gamma)
    kw = {"precision": precision}
    if device is not None:
        kw["device"] = device
    g = pr.pyrandaGroupSim("KH-wall-group", [_kh_wall_mesh(nx, ny) for nx, ny in sizes], **kw)
    g.addPackage(pr.pyrandaTimestep(g))
    g.addPackage(pr.pyrandaBC(g))
    g.EOM(KH_EOM + KH_WALL_BC, {"gamma": gamma, "R0": R, "cv": cp - R})
    g.setIC(kh_wall_ic(seed))
    dt = g.var("dt").data
    return g, {"time": torch.zeros(len(sizes), dtype=torch.float64, device=dt.device), "dt": dt.clone()}
