"""Command-line front-end for hard-wall contact calculations (``ce_hard_wall``).

Mirrors ``ContactMechanics/CommandLineInterface/HardWall.py:36-558``: the same
options, the same three load programmes (a pressure range or a pressure file,
a displacement range, or by default ``NSTEPS`` displacements chosen so that
the contact areas spread evenly on a log scale), the same text outputs
(contact, pressure, displacement and gap maps with a header of macroscopic
quantities; a load/area log) and NetCDF frames (``displacements``,
``forces``, ``displacement``, ``load``, ``area``).  Every solve runs on the
GPU through ``NonSmoothContactSystem.minimize_proxy``.

Run as ``python -m contactmechanics_amd.CommandLineInterface.HardWall FILE
[options]``.  Topography files: a text height matrix (``numpy.loadtxt``),
``.npy``, or NetCDF with a 2D ``heights`` variable; SurfaceTopography's other
readers are not available here.
"""

import sys
from argparse import ArgumentParser, ArgumentTypeError

import numpy as np

from .. import FreeFFTElasticHalfSpace, PeriodicFFTElasticHalfSpace, __version__, make_plastic_system, make_system
from ..IO.NetCDF import NetCDFContainer
from ..Tools.Logger import Logger, quiet, screen
from ..Topography import PlasticTopography, Topography

NSTEPS = 20  # displacements of the automatic programme

UNIT_TO_METERS = {"A": 1e-10, "nm": 1e-9, "µm": 1e-6, "mm": 1e-3, "m": 1.0, "unknown": 1.0}

VERSION = "contactmechanics_amd version: {}".format(__version__)


def read_topography(fn, physical_sizes=None, unit=None):
    """A ``Topography`` from a text matrix, ``.npy`` or a NetCDF file with a
    2D ``heights`` variable (``length_x``/``length_y`` attributes, if
    present, give the physical sizes unless ``physical_sizes`` does)."""
    if fn.endswith(".npy"):
        h = np.load(fn, allow_pickle=False)
    elif fn.endswith(".nc"):
        from scipy.io import netcdf_file
        with netcdf_file(fn, "r", mmap=False) as nc:
            h = np.array(nc.variables["heights"][...], dtype=float)
            attrs = nc._attributes
            if physical_sizes is None and "length_x" in attrs and "length_y" in attrs:
                physical_sizes = (float(attrs["length_x"]), float(attrs["length_y"]))
            if unit is None and "unit" in attrs:
                u = attrs["unit"]
                unit = u.decode() if isinstance(u, bytes) else u
    else:
        h = np.loadtxt(fn)
    return Topography(np.atleast_2d(h), physical_sizes, unit=unit)


def next_step(system, surface, history=None, pentol=None, maxiter=None, logger=quiet, verbose=False):
    """One solve of the automatic programme (reference :40-164): the first
    displacement is minus the mean height, the second just below the top, and
    each later one bisects the widest gap between neighbouring contact areas
    on a log scale (or, when the widest gap is above the largest area,
    doubles the distance of the deepest displacement from the bottom).
    Returns the contact set, displacements and forces on the topography, the
    displacement, load (mean pressure) and fractional area, and the updated
    history (displacements, mean gaps, loads, areas, convergence flags, sorted
    by area)."""
    h = np.asarray(surface.heights())
    top, middle, bot = float(np.max(h)), float(np.mean(h)), float(np.min(h))
    if history is None:
        disp, gap, load, area, converged = [], [], [], [], np.array([], dtype=bool)
    else:
        disp, gap, load, area, converged = (list(history[0]), list(history[1]), list(history[2]),
                                            list(history[3]), np.asarray(history[4], dtype=bool))
    step = len(disp)
    if step == 0:
        disp0 = -middle
    elif step == 1:
        disp0 = -top + 0.01 * (top - middle)
    else:
        log_area = np.log10(np.array(area) + 1 / np.prod(surface.nb_grid_pts))
        widths = np.append(np.diff(log_area), -log_area[-1])
        i = int(np.argmax(widths))
        disp0 = bot + 2 * (disp[-1] - bot) if i == step - 1 else (disp[i] + disp[i + 1]) / 2
    opt = system.minimize_proxy(offset=disp0, logger=logger, pentol=pentol, maxiter=maxiter, verbose=verbose)
    f = opt.jac
    u = opt.x[:f.shape[0], :f.shape[1]]
    cur_load = f.sum() / np.prod(surface.physical_sizes)
    cur_area = (f > 0).sum() / np.prod(surface.nb_grid_pts)
    rows = sorted(zip(disp + [disp0], gap + [np.mean(u) - middle - disp0], load + [cur_load], area + [cur_area],
                      list(converged) + [bool(opt.success)]), key=lambda r: r[3])
    logger.pr("disp = {}, area = {}, load = {}, converged = {}".format(disp0, cur_area, cur_load, opt.success))
    cols = [np.array([r[k] for r in rows], dtype=float) for k in range(4)]
    history = (cols[0], cols[1], cols[2], cols[3], np.array([r[4] for r in rows], dtype=bool))
    return opt.active_set, u, f, disp0, cur_load, cur_area, history


def dump(txt, surface, substrate, u, f, offset=0):
    """Write one row of macroscopic quantities to ``txt`` and return them as
    (header, value) pairs (reference :167-195)."""
    unit = surface.unit
    mean_elastic = float(np.mean(u))
    mean_rigid = float(np.mean(surface[...])) + offset
    load = float(f.sum())
    area_px = int((f > 0).sum())
    e_unit = "E*" if substrate.young == 1 else "[Units of E*]"
    header = ["mean elastic ({})".format(unit), "mean rigid ({})".format(unit), "mean gap ({})".format(unit),
              "load ({} {}^2)".format(e_unit, unit),
              "mean pressure ({})".format(e_unit), "area ({}^2)".format(unit), "fractional area"]
    data = [mean_elastic, mean_rigid, mean_elastic - mean_rigid, load, load / np.prod(surface.physical_sizes),
            area_px * surface.area_per_pt, area_px / np.prod(surface.nb_grid_pts)]
    txt.st(header, data)
    return list(zip(header, data))


def dump_nc(container, u, f, disp0, load, area):
    if container is not None:
        frame = container.get_next_frame()
        frame.displacements = np.asarray(u, dtype=float)
        frame.forces = np.asarray(f, dtype=float)
        frame.displacement = float(disp0)
        frame.load = float(load)
        frame.area = float(area)


def _header(macro, tail):
    lines = [VERSION, " ".join(sys.argv)]
    if macro:
        lines.append("\n".join("{} = {}".format(k, v) for k, v in macro))
    return "\n".join(lines) + tail


def save_contact(fn, surface, substrate, contact, macro=None):
    np.savetxt(fn, contact, fmt="%i", header=_header(macro, "Contact map follows. Values are boolean."))


def save_pressure(fn, surface, substrate, pressure, macro=None):
    unit = ("Pressure values follow, they are reported in units of E*." if substrate.young == 1 else
            "This calculation was run with a contact modulus E*={}.".format(substrate.young))
    np.savetxt(fn, pressure, header=_header(macro, unit))


def save_gap(fn, surface, gap, macro=None):
    unit = ("No unit information available." if surface.unit is None else
            "Gap values follow, they are reported in units of {}.".format(surface.unit))
    np.savetxt(fn, gap, header=_header(macro, unit))


def tuple2(s):
    try:
        x, y = (float(v) for v in s.split(","))
    except (ValueError, TypeError):
        raise ArgumentTypeError("Size must be sx,sy")
    return x, y


def _parser():
    p = ArgumentParser(description="Run a contact mechanics calculation with a hard-wall interaction using "
                                   "Polonsky & Keer's constrained conjugate gradient solver.")
    p.add_argument("filename", metavar="FILENAME", help="name of topography file")
    p.add_argument("--detrend", dest="detrend", type=str, metavar="DETREND",
                   help='detrend surface before contact calculation: "height", "slope" or "curvature"')
    p.add_argument("--boundary", dest="boundary", type=str, default="periodic", metavar="BOUNDARY",
                   help="boundary conditions, BOUNDARY=periodic|nonperiodic")
    p.add_argument("--modulus", dest="modulus", type=float, default=1.0, metavar="MODULUS",
                   help="contact/Young's modulus")
    p.add_argument("--poisson", dest="poisson", type=float, default=0.0, metavar="POISSON", help="Poisson number")
    p.add_argument("--thickness", dest="thickness", type=float, default=None, metavar="THICKNESS",
                   help="substrate of thickness THICKNESS on a rigid base")
    p.add_argument("--hardness", dest="hardness", type=float, metavar="HARDNESS", help="penetration hardness")
    p.add_argument("--maxiter", dest="maxiter", type=int, default=1000, metavar="MAXITER",
                   help="stop after MAXITER iterations")
    p.add_argument("--displacement", dest="displacement", type=str, metavar="DISPLACEMENT",
                   help="displacement, or range min,max,steps")
    p.add_argument("--pressure", dest="pressure", type=str, metavar="PRESSURE",
                   help="external pressure, or range min,max,steps")
    p.add_argument("--pressure-from-file", dest="pressure_from_fn", type=str, metavar="PRESSUREFN",
                   help="external pressures from the file PRESSUREFN")
    p.add_argument("--physical_sizes", dest="physical_sizes", type=tuple2, metavar="SIZE",
                   help="physical sizes sx,sy")
    p.add_argument("--physical_sizes-unit", dest="size_unit", type=str, metavar="UNIT", help="length unit")
    p.add_argument("--height-fac", dest="height_fac", type=float, metavar="FAC", help="scale all heights by FAC")
    p.add_argument("--height-unit", dest="height_unit", type=str, metavar="UNIT", help="height unit")
    p.add_argument("--pentol", dest="pentol", type=float, metavar="PENTOL", help="penetration tolerance")
    p.add_argument("--contact-fn", dest="contact_fn", type=str, metavar="CONTACTFN", help="contact map file")
    p.add_argument("--pressure-fn", dest="pressure_fn", type=str, metavar="PRESSUREFN", help="pressure map file")
    p.add_argument("--displ-fn", dest="displ_fn", type=str, metavar="DISPLFN", help="displacement map file")
    p.add_argument("--gap-fn", dest="gap_fn", type=str, metavar="GAPFN", help="gap map file")
    p.add_argument("--log-fn", dest="log_fn", type=str, default="hard_wall.out", metavar="LOGFN",
                   help="log file of final area and load")
    p.add_argument("--netcdf-fn", dest="netcdf_fn", type=str, default=None, metavar="NETCDFFN",
                   help="NetCDF file of displacement and force frames")
    p.add_argument("--verbose", dest="verbose", action="store_true", default=False, help="verbose output")
    return p


def _range(spec, what):
    v = spec.split(",")
    if len(v) == 1:
        return [float(v[0])]
    if len(v) == 3:
        return list(np.linspace(float(v[0]), float(v[1]), int(v[2])))
    print("Please specify either single {0} value or 3-tuple for {0} range.".format(what))
    sys.exit(999)


def main(argv=None, logger=screen):
    args = _parser().parse_args(argv)
    logger.pr(VERSION)
    for k, v in vars(args).items():
        logger.pr("{} = {}".format(k, v))

    unit = args.size_unit
    surface = read_topography(args.filename, args.physical_sizes, unit)
    if surface.unit is None:
        surface.unit = "N/A"
    if args.height_fac is not None or args.height_unit is not None:
        fac = args.height_fac if args.height_fac is not None else 1.0
        if args.height_unit is not None:
            fac *= UNIT_TO_METERS[args.height_unit] / UNIT_TO_METERS.get(surface.unit, 1.0)
        logger.pr("Rescaling surface heights by {}.".format(fac))
        surface = surface.scale(fac)
    logger.pr("SurfaceTopography has dimension of {} and physical_sizes of {} {}.".format(
        surface.nb_grid_pts, surface.physical_sizes, surface.unit))
    logger.pr("RMS height = {}, RMS gradient = {}".format(surface.rms_height_from_area(), surface.rms_gradient()))
    if args.detrend is not None:
        surface = surface.detrend(detrend_mode=args.detrend)
        logger.pr("After detrending: RMS height = {}, RMS gradient = {}".format(
            surface.rms_height_from_area(), surface.rms_gradient()))
    if args.hardness is not None:
        surface = PlasticTopography(surface, args.hardness)

    if args.boundary == "periodic":
        substrate = PeriodicFFTElasticHalfSpace(surface.nb_grid_pts, args.modulus, surface.physical_sizes,
                                                thickness=args.thickness, poisson=args.poisson)
    elif args.boundary == "nonperiodic":
        if args.thickness is not None:
            raise ValueError('"thickness" arguments cannot be used with nonperiodic boundaries.')
        substrate = FreeFFTElasticHalfSpace(surface.nb_grid_pts, args.modulus / (1 - args.poisson ** 2),
                                            surface.physical_sizes)
    else:
        raise ValueError("Unknown boundary conditions: {}".format(args.boundary))
    system = (make_plastic_system if args.hardness is not None else make_system)(substrate=substrate, surface=surface)

    container = None
    if args.netcdf_fn is not None:
        container = NetCDFContainer(args.netcdf_fn, mode="w", double=True)
        container.set_shape(surface.nb_grid_pts)
    txt = Logger(args.log_fn)

    def outputs(i, n, contact, u, f, offset):
        suffix = "" if n == 1 else ".{}".format(i)
        macro = dump(txt, surface, substrate, u, f, offset)
        if args.contact_fn is not None:
            save_contact(args.contact_fn + suffix, surface, substrate, contact, macro=macro)
        if args.pressure_fn is not None:
            save_pressure(args.pressure_fn + suffix, surface, substrate, f / surface.area_per_pt, macro=macro)
        if args.displ_fn is not None:
            save_gap(args.displ_fn + suffix, surface, u, macro=macro)
        if args.gap_fn is not None:
            save_gap(args.gap_fn + suffix, surface, u - surface[...] - offset, macro=macro)

    def report(opt, f):
        logger.pr("displacement = {}".format(opt.offset))
        logger.pr("pressure = {}".format(f.sum() / np.prod(surface.physical_sizes)))
        logger.pr("energy = {}".format(opt.fun))
        logger.pr("fractional contact area = {}".format((f > 0).sum() / np.prod(surface.nb_grid_pts)))

    if args.pressure is not None or args.pressure_from_fn is not None:
        if args.displacement is not None:
            raise ValueError("Please specify either displacement or pressure range, not both.")
        pressures = (_range(args.pressure, "pressure") if args.pressure is not None
                     else list(np.ravel(np.loadtxt(args.pressure_from_fn))))
        for i, p in enumerate(pressures):
            external_force = p * np.prod(surface.physical_sizes)
            logger.pr("external_force = {}".format(external_force))
            opt = system.minimize_proxy(logger=logger, external_force=external_force, pentol=args.pentol,
                                        maxiter=args.maxiter, verbose=args.verbose)
            f = opt.jac
            u = opt.x[:f.shape[0], :f.shape[1]]
            report(opt, f)
            dump_nc(container, u, f, opt.offset, p, (f > 0).sum() / np.prod(surface.nb_grid_pts))
            outputs(i, len(pressures), opt.active_set, u, f, opt.offset)
    elif args.displacement is not None:
        displacements = _range(args.displacement, "displacement")
        for i, d in enumerate(displacements):
            opt = system.minimize_proxy(offset=d, logger=logger, pentol=args.pentol, maxiter=args.maxiter,
                                        verbose=args.verbose)
            f = opt.jac
            u = opt.x[:f.shape[0], :f.shape[1]]
            report(opt, f)
            dump_nc(container, u, f, opt.offset, f.sum() / np.prod(surface.physical_sizes),
                    (f > 0).sum() / np.prod(surface.nb_grid_pts))
            outputs(i, len(displacements), opt.active_set, u, f, opt.offset)
    else:
        history = None
        for i in range(NSTEPS):
            contact, u, f, disp0, load, area, history = next_step(
                system, surface, history, pentol=args.pentol, maxiter=args.maxiter, logger=logger,
                verbose=args.verbose)
            dump_nc(container, u, f, disp0, load, area)
            outputs(i, NSTEPS, contact, u, f, disp0)
    if container is not None:
        container.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
