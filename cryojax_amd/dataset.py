"""RELION particle datasets: STAR parameter files, MRCS particle stacks and
simulate_particle_stack, whose images come from the device forward.

Mirrors cryojax.dataset (cryojax/dataset/_particle_data/relion.py:35-1710:
RelionParticleParameterFile, RelionParticleStackDataset and their STAR <->
parameter conversions; particle_simulation.py:22-262 simulate_particle_stack).
The conversions keep RELION's conventions as the reference states them:

* pose: EulerAnglePose(-rlnOriginXAngst, -rlnOriginYAngst, -rlnAngleRot,
  -rlnAngleTilt, -rlnAnglePsi) (relion.py:1053-1104, 1147-1153), psi -999
  (RELION's "not estimated") read as 0 (:1070-1093), tilt / psi priors used
  when the angles are absent (:1064-1097);
* CTF: defocus = (U + V) / 2, astigmatism = U - V, angle = rlnDefocusAngle,
  phase shift, with Cs and amplitude contrast from the optics group
  (:986-1016); written back as U = defocus + z + astig / 2,
  V = defocus + z - astig / 2 (:1510-1527);
* optics group: image size, pixel size, voltage, Cs, amplitude contrast,
  appended as a new group per write (:1384-1445, 1572-1577);
* image names: "<i+1 padded to n_characters>@<output_folder>/<prefix><delim>
  <file number>.mrcs", file numbers continuing from the previous row's file
  (:1630-1689).

* CTF envelope (loads_envelope=True): FourierGaussian(rlnCtfScalefactor,
  rlnCtfBfactor), or Constant(rlnCtfScalefactor) without a B-factor, per
  particle (:1016-1029, 1156-1176); written back from a FourierGaussian or a
  Constant envelope (:1526-1540).
"""

from __future__ import annotations

import pathlib
import re
from typing import Any, Callable

import numpy as np
import pandas as pd

import warnings

from . import io
from .simulator import AstigmaticCTF, BasicImageConfig, ContrastTransferTheory, EulerAnglePose
from .transforms import Constant, FourierGaussian

RELION_CTF_OPTICS_ENTRIES = [("rlnSphericalAberration", "Float64"), ("rlnAmplitudeContrast", "Float64")]
RELION_INSTRUMENT_OPTICS_ENTRIES = [("rlnImageSize", "Int64"), ("rlnVoltage", "Float64"),
                                    ("rlnImagePixelSize", "Float64")]
RELION_CTF_PARTICLE_ENTRIES = [("rlnDefocusU", "Float64"), ("rlnDefocusV", "Float64"), ("rlnDefocusAngle", "Float64"),
                               ("rlnPhaseShift", "Float64")]
RELION_POSE_PARTICLE_ENTRIES = [("rlnOriginXAngst", "Float64"), ("rlnOriginYAngst", "Float64"),
                                ("rlnAngleRot", "Float64"), ("rlnAngleTilt", "Float64"), ("rlnAnglePsi", "Float64")]
RELION_DEFAULT_OPTICS_ENTRIES = [*RELION_INSTRUMENT_OPTICS_ENTRIES, *RELION_CTF_OPTICS_ENTRIES,
                                 ("rlnOpticsGroup", "Int64")]
RELION_DEFAULT_PARTICLE_ENTRIES = [*RELION_CTF_PARTICLE_ENTRIES, *RELION_POSE_PARTICLE_ENTRIES,
                                   ("rlnOpticsGroup", "Int64")]
RELION_REQUIRED_OPTICS_ENTRIES = RELION_DEFAULT_OPTICS_ENTRIES
RELION_REQUIRED_PARTICLE_ENTRIES = [*RELION_CTF_PARTICLE_ENTRIES, ("rlnOpticsGroup", "Int64")]
RELION_SUPPORTED_PARTICLE_ENTRIES = [*RELION_REQUIRED_PARTICLE_ENTRIES, ("rlnCtfBfactor", "Float64"),
                                     ("rlnCtfScalefactor", "Float64")]


def _validate_mode(mode):
    if mode not in ("r", "w"):
        raise ValueError(f"Passed unsupported `mode = {mode}`. Supported modes are 'r' and 'w'.")
    return mode


def _empty_frame(entries):
    return pd.DataFrame({c: pd.Series(dtype=d) for c, d in entries})


def _validate_starfile_data(data):
    """relion.py:1295-1315."""
    for block, entries in (("particles", RELION_REQUIRED_PARTICLE_ENTRIES), ("optics", RELION_REQUIRED_OPTICS_ENTRIES)):
        if block not in data:
            raise ValueError(f"Missing key '{block}' in the STAR file.")
        need = [c for c, _ in entries]
        if not set(need).issubset(data[block].columns):
            raise ValueError(f"Missing required keys in starfile '{block}' group. Required keys are {tuple(need)}.")


def _select_particles(data, selection_filter):
    """relion.py:926-975."""
    particles = data["particles"]
    mask = pd.Series(True, index=particles.index)
    for key, fn in selection_filter.items():
        if key not in particles.columns:
            raise ValueError(f"Included key '{key}' in the `selection_filter`, but this entry could not be found in "
                             "the STAR file.")
        if not callable(fn):
            raise ValueError(f"`selection_filter['{key}']` must be a function returning a boolean mask.")
        m = fn(particles[key])
        if not pd.api.types.is_bool_dtype(m):
            raise ValueError(f"`selection_filter['{key}']` did not return a boolean mask.")
        mask = mask & m
    data["particles"] = particles[mask]
    return data


def _load_starfile_data(path, selection_filter, mode, exists_ok):
    """relion.py:865-915."""
    if mode == "r":
        if not path.exists():
            raise FileNotFoundError(f"Set `mode = 'r'`, but STAR file {path} does not exist. To write a new STAR "
                                    "file, set `mode = 'w'`.")
        data = io.read_starfile(path)
        _validate_starfile_data(data)
        if selection_filter is not None:
            data = _select_particles(data, selection_filter)
        return {"optics": data["optics"], "particles": data["particles"]}
    if path.exists() and not exists_ok:
        raise FileExistsError(f"Set `mode = 'w'`, but STAR file {path} already exists. To read it set `mode = 'r'`, "
                              "or set `exists_ok=True` to replace it.")
    if selection_filter is not None:
        raise ValueError("A `selection_filter` is only used in `mode = 'r'`.")
    return {"optics": _empty_frame(RELION_DEFAULT_OPTICS_ENTRIES),
            "particles": _empty_frame(RELION_DEFAULT_PARTICLE_ENTRIES)}


def _validate_dataset_index(cls, index, n_rows):
    """relion.py:1264-1292."""
    msg = (f"The index at which the `{cls.__name__}` was accessed was out of bounds! The number of rows in the "
           f"dataset is {n_rows}, but you tried to access the index {{}}.")
    if isinstance(index, (int, np.integer)):
        if index > n_rows - 1:
            raise IndexError(msg.format(index))
    elif isinstance(index, slice):
        if index.start is not None and index.start > n_rows - 1:
            raise IndexError(msg.format(index.start))
    elif isinstance(index, np.ndarray):
        if index.size == 0:
            raise IndexError("Found that the index passed to the dataset was an empty numpy array.")
    else:
        raise IndexError(f"Indexing with the type {type(index)} is not supported by `{cls.__name__}`.")


def _optics_group_of(particles_at_index, optics):
    """relion.py:1580-1610 (one optics group per read)."""
    groups = np.unique(np.atleast_1d(np.asarray(particles_at_index["rlnOpticsGroup"])))
    if groups.size > 1:
        raise NotImplementedError("Tried to read multiple particles at once that belong to different optics groups.")
    return optics[optics["rlnOpticsGroup"] == int(groups[0])].iloc[0]


def _column(particles, key, default=0.0):
    return np.asarray(particles[key], dtype=np.float64) if key in particles else np.asarray(default, dtype=np.float64)


def _invert_rotation(pose):
    """relion.py:1198-1208: (phi, theta, psi) -> (-psi, -theta, -phi), wrapped to [-180, 180)."""
    neg = lambda a: ((-np.asarray(a) + 180.0) % 360.0) - 180.0  # noqa: E731
    return EulerAnglePose(pose.offset_in_angstroms[..., 0], pose.offset_in_angstroms[..., 1],
                          neg(pose.psi_angle), neg(pose.theta_angle), neg(pose.phi_angle),
                          offset_z_in_angstroms=pose.offset_z_in_angstroms)


def _make_pytrees_from_starfile(particles, optics, broadcasts_image_config, loads_envelope, pad_options,
                                inverts_rotation):
    """relion.py:978-1131."""
    single = isinstance(particles, pd.Series)
    du, dv = _column(particles, "rlnDefocusU"), _column(particles, "rlnDefocusV")
    defocus, astig = (du + dv) / 2, du - dv
    angle, phase = _column(particles, "rlnDefocusAngle"), _column(particles, "rlnPhaseShift")
    batch = () if single else (len(particles),)
    cs = np.full(batch, float(optics["rlnSphericalAberration"]))
    ac = np.full(batch, float(optics["rlnAmplitudeContrast"]))
    envelope = None
    if loads_envelope:
        keys_ = particles.index if single else particles.columns
        b = np.asarray(particles["rlnCtfBfactor"], dtype=np.float64) if "rlnCtfBfactor" in keys_ else None
        amp = np.asarray(particles["rlnCtfScalefactor"], dtype=np.float64) if "rlnCtfScalefactor" in keys_ else None
        envelope = _make_envelope_function(amp, b)
    pixel_size = np.asarray(float(optics["rlnImagePixelSize"]))
    voltage = np.asarray(float(optics["rlnVoltage"]))
    if broadcasts_image_config and batch:
        pixel_size, voltage = np.full(batch, pixel_size), np.full(batch, voltage)
    keys = particles.index if single else particles.columns
    ox = particles["rlnOriginXAngst"] if "rlnOriginXAngst" in keys else 0.0
    oy = particles["rlnOriginYAngst"] if "rlnOriginYAngst" in keys else 0.0
    rot = particles["rlnAngleRot"] if "rlnAngleRot" in keys else 0.0
    if "rlnAngleTilt" in keys:
        tilt = particles["rlnAngleTilt"]
    elif "rlnAngleTiltPrior" in keys:
        tilt = particles["rlnAngleTiltPrior"]
    else:
        tilt = 0.0
    if "rlnAnglePsi" in keys:
        psi = np.asarray(particles["rlnAnglePsi"], dtype=np.float64)
        psi = np.where(psi == -999.0, 0.0, psi)
    elif "rlnAnglePsiPrior" in keys:
        psi = particles["rlnAnglePsiPrior"]
    else:
        psi = 0.0
    full = lambda v: np.full(batch, np.asarray(v, dtype=np.float64)) if batch and np.ndim(v) == 0 else \
        np.asarray(v, dtype=np.float64)  # noqa: E731
    pose = EulerAnglePose(-full(ox), -full(oy), -full(rot), -full(tilt), -full(psi))
    if inverts_rotation:
        pose = _invert_rotation(pose)
    size = int(optics["rlnImageSize"])
    config = BasicImageConfig((size, size), pixel_size, voltage, pad_options=pad_options)
    theory = ContrastTransferTheory(AstigmaticCTF(defocus, astig, angle, cs), envelope=envelope,
                                    amplitude_contrast_ratio=ac, phase_shift=phase)
    return config, theory, pose


def _make_envelope_function(amp, b_factor):
    """relion.py:1156-1176: no columns -> a warning and no envelope; a scale factor alone -> a
    Constant; otherwise a FourierGaussian, with unit amplitude when the scale factor is absent."""
    if b_factor is None and amp is None:
        warnings.warn("`loads_envelope` was set to True, but no envelope parameters were found. Setting envelope "
                      "as None. Make sure your starfile is correctly formatted or set `loads_envelope=False`.")
        return None
    # the reference swaps the arrays into default-built operators (eqx.tree_at), which skips the
    # constructors' checks: a B-factor of zero, common in RELION files, is kept
    if b_factor is None:
        env = Constant(1.0)
        env.value = amp
        return env
    env = FourierGaussian(1.0, 1.0)
    env.amplitude, env.b_factor = (np.ones_like(b_factor) if amp is None else amp), b_factor
    return env


def _parameters_to_optics_data(parameters, group):
    """relion.py:1384-1444."""
    if not {"image_config", "transfer_theory"}.issubset(parameters):
        raise ValueError("Tried to add an optics group, but the parameters have no 'image_config' and "
                         "'transfer_theory'.")
    config, theory = parameters["image_config"], parameters["transfer_theory"]
    if config.shape[0] != config.shape[1]:
        raise ValueError("Only square image shapes can be written to an optics group.")
    if not isinstance(theory.ctf, AstigmaticCTF):
        raise ValueError("`transfer_theory.ctf` must be an `AstigmaticCTF`.")
    entries = {"rlnOpticsGroup": group, "rlnImageSize": config.shape[0], "rlnImagePixelSize": config.pixel_size,
               "rlnVoltage": config.voltage_in_kilovolts, "rlnSphericalAberration": theory.ctf.spherical_aberration_in_mm,
               "rlnAmplitudeContrast": theory.amplitude_contrast_ratio}
    row = {}
    for k, v in entries.items():
        a = np.atleast_1d(np.asarray(v))
        if a.size > 1 and np.unique(a).size > 1:
            raise ValueError(f"Optics group entry '{k}' must be a scalar or a constant array; got {v}.")
        row[k] = [a.ravel()[0]]
    return pd.DataFrame(row)


def _parameters_to_particle_data(parameters, group=None):
    """relion.py:1447-1569."""
    if "pose" not in parameters:
        raise ValueError("The parameters have no 'pose' entry.")
    pose = parameters["pose"]
    if not isinstance(pose, EulerAnglePose):
        raise ValueError("`pose` must be an `EulerAnglePose`.")
    off = np.asarray(pose.offset_in_angstroms, dtype=np.float64)
    cols = {"rlnOriginXAngst": -off[..., 0], "rlnOriginYAngst": -off[..., 1], "rlnAngleRot": -pose.phi_angle,
            "rlnAngleTilt": -pose.theta_angle, "rlnAnglePsi": -pose.psi_angle}
    n = int(np.asarray(off[..., 0]).size)
    if "transfer_theory" in parameters:
        theory = parameters["transfer_theory"]
        if not isinstance(theory.ctf, AstigmaticCTF):
            raise ValueError("`transfer_theory.ctf` must be an `AstigmaticCTF`.")
        z = 0.0 if pose.offset_z_in_angstroms is None else pose.offset_z_in_angstroms
        ctf = theory.ctf
        cols["rlnDefocusU"] = ctf.defocus_in_angstroms + z + ctf.astigmatism_in_angstroms / 2
        cols["rlnDefocusV"] = ctf.defocus_in_angstroms + z - ctf.astigmatism_in_angstroms / 2
        cols["rlnDefocusAngle"] = ctf.astigmatism_angle
        env = theory.envelope
        if isinstance(env, FourierGaussian):
            cols["rlnCtfBfactor"] = env.b_factor
            cols["rlnCtfScalefactor"] = env.amplitude
        elif isinstance(env, Constant):
            cols["rlnCtfScalefactor"] = env.value
        elif env is not None:
            raise ValueError("When adding particle to STAR file, `transfer_theory.envelope` must be type "
                             "`FourierGaussian` or `Constant`, or `None`. Got "
                             f"{type(env).__name__}.")
        cols["rlnPhaseShift"] = theory.phase_shift
    for k, v in cols.items():
        a = np.asarray(v, dtype=np.float64)
        if a.ndim == 0:
            cols[k] = np.full(n, a)
        elif a.size == n:
            cols[k] = a.ravel()
        else:
            raise ValueError("Found an inconsistent number of particles in the parameters.")
    if group is not None:
        cols["rlnOpticsGroup"] = np.full(n, group, dtype=np.int64)
    frame = pd.DataFrame(cols)
    meta = parameters.get("metadata")
    if meta is not None:
        if len(meta) != n:
            raise ValueError("The metadata's particle count differs from the pose's.")
        frame = pd.concat([frame, meta.reset_index(drop=True)], axis="columns", verify_integrity=True)
    return frame


def _validate_parameters(parameters, force_keys=False):
    """relion.py:1351-1381."""
    if force_keys and not {"image_config", "transfer_theory", "pose"}.issubset(parameters):
        raise ValueError("The parameters must have keys 'pose', 'transfer_theory' and 'image_config'.")
    for key, cls in (("image_config", BasicImageConfig), ("transfer_theory", ContrastTransferTheory),
                     ("pose", EulerAnglePose)):
        if key in parameters and not isinstance(parameters[key], cls):
            raise TypeError(f"Found that dict key '{key}' was not type `{cls.__name__}`.")


def _next_optics_group(optics):
    groups = np.asarray(optics["rlnOpticsGroup"], dtype=np.int64)
    return (0 if groups.size == 0 else int(groups[-1])) + 1


class RelionParticleParameterFile:
    """relion.py:204-533."""

    def __init__(self, path_to_starfile, mode="r", exists_ok=False, selection_filter=None, *, loads_metadata=False,
                 broadcasts_image_config=True, loads_envelope=False, updates_optics_group=False,
                 inverts_rotation=False, pad_options=None):
        self._pad_options = dict(pad_options or {})
        self._mode = _validate_mode(mode)
        self._path_to_starfile = pathlib.Path(path_to_starfile)
        self._starfile_data = _load_starfile_data(self._path_to_starfile, selection_filter, mode, exists_ok)
        self.loads_metadata = loads_metadata
        self.broadcasts_image_config = broadcasts_image_config
        self.loads_envelope = loads_envelope
        self.updates_optics_group = updates_optics_group
        self.inverts_rotation = inverts_rotation

    @property
    def mode(self):
        return self._mode

    @property
    def path_to_starfile(self):
        return self._path_to_starfile

    @path_to_starfile.setter
    def path_to_starfile(self, value):
        self._path_to_starfile = pathlib.Path(value)

    @property
    def starfile_data(self):
        return self._starfile_data

    @starfile_data.setter
    def starfile_data(self, value):
        _validate_starfile_data(value)
        self._starfile_data = {"optics": value["optics"], "particles": value["particles"]}

    def __len__(self):
        return len(self._starfile_data["particles"])

    def __getitem__(self, index):
        particles = self._starfile_data["particles"]
        _validate_dataset_index(type(self), index, len(particles))
        at = particles.iloc[index]
        optics = _optics_group_of(at, self._starfile_data["optics"])
        config, theory, pose = _make_pytrees_from_starfile(at, optics, self.broadcasts_image_config,
                                                           self.loads_envelope, self._pad_options,
                                                           self.inverts_rotation)
        metadata = None
        if self.loads_metadata:
            frame = at.to_frame().T if isinstance(at, pd.Series) else at
            known = {c for c, _ in RELION_SUPPORTED_PARTICLE_ENTRIES}
            metadata = frame.drop([c for c in frame.columns if c in known], axis="columns")
        return dict(image_config=config, pose=pose, transfer_theory=theory, metadata=metadata)

    def __setitem__(self, index, value):
        particles = self._starfile_data["particles"]
        _validate_dataset_index(type(self), index, len(particles))
        _validate_parameters(value)
        value = dict(value)
        if "pose" in value and self.inverts_rotation:
            value["pose"] = _invert_rotation(value["pose"])
        optics = self._starfile_data["optics"]
        if self.updates_optics_group:
            group = _next_optics_group(optics)
            update = _parameters_to_particle_data(value, group)
            optics = pd.concat([optics, _parameters_to_optics_data(value, group)], ignore_index=True)
        else:
            update = _parameters_to_particle_data(value)
        for c in set(update.columns) - set(particles.columns):
            particles[c] = pd.Series(dtype=update[c].dtype)
        rows = np.atleast_1d(np.arange(len(particles))[index])
        particles.loc[particles.index[rows], update.columns] = update.values
        self._starfile_data = {"optics": optics, "particles": particles}

    def append(self, value):
        """relion.py:382-418."""
        _validate_parameters(value, force_keys=True)
        value = dict(value)
        if self.inverts_rotation:
            value["pose"] = _invert_rotation(value["pose"])
        optics, particles = self._starfile_data["optics"], self._starfile_data["particles"]
        group = _next_optics_group(optics)
        new_optics = _parameters_to_optics_data(value, group)
        new_particles = _parameters_to_particle_data(value, group)
        optics = pd.concat([optics, new_optics], ignore_index=True) if len(optics) else new_optics
        particles = pd.concat([particles, new_particles], ignore_index=True) if len(particles) else new_particles
        self._starfile_data = {"optics": optics, "particles": particles}

    def save(self, *, overwrite=False, **kwargs):
        """relion.py:421-438."""
        path = self._path_to_starfile
        if path.exists() and not overwrite:
            raise FileExistsError(f"Tried saving STAR file, but file {path} already exists. Set overwrite=True.")
        path.parent.mkdir(parents=True, exist_ok=True)
        io.write_starfile(self._starfile_data, path, **kwargs)

    def copy(self):
        out = RelionParticleParameterFile.__new__(RelionParticleParameterFile)
        out.__dict__.update(self.__dict__)
        out._starfile_data = {k: v.copy() for k, v in self._starfile_data.items()}
        return out


def _mrcfile_settings(d):
    """relion.py:1613-1627."""
    d = dict(d or {})
    return dict(prefix=d.get("prefix", ""), output_folder=d.get("output_folder", ""), delimiter=d.get("delimiter", "_"),
                n_characters=d.get("n_characters", 6), overwrite=d.get("overwrite", False),
                compression=d.get("compression", None))


def _format_number(number, n_characters=6):
    """relion.py:1702-1707."""
    return str(int(number)).zfill(n_characters) if number else "0" * n_characters


def _file_number_of(filename):
    """relion.py:1690-1699: the number before the extension."""
    m = re.search(r"(\d+)\.[^.]+$", filename)
    if m is None:
        raise OSError(f"Could not get the file number from file {filename}.")
    return int(m.group(1))


def _make_image_filename(index, particles, n_particles, settings, project):
    """relion.py:1636-1687."""
    file_number = 0
    if n_particles and index[0] - 1 >= 0:
        last = particles["rlnImageName"].iloc[index[0] - 1]
        if not isinstance(last, str) or "@" not in last:
            raise OSError(f"Could not read the previous file number at index {int(index[0] - 1)}.")
        file_number = _file_number_of(last.split("@")[1]) + 1
    number = _format_number(file_number, settings["n_characters"])
    name = number + ".mrcs" if settings["prefix"] == "" else settings["prefix"] + settings["delimiter"] + number + ".mrcs"
    rel = str(pathlib.Path(settings["output_folder"], name))
    names = [_format_number(i + 1, settings["n_characters"]) + "@" + rel for i in range(index.size)]
    return pathlib.Path(project, rel), names


def _load_image_stack_from_mrc(shape, particles, project):
    """relion.py:1211-1261."""
    names = particles["rlnImageName"].reset_index(drop=True)
    split = names.astype(str).str.split("@", n=1)
    if split.str.len().min() != 2:
        raise ValueError("Could not parse filenames from the 'rlnImageName' column (expected 'index@file.mrcs').")
    mrc_index = split.str[0].astype(int) - 1
    files = split.str[1]
    stack = np.empty((len(files), *shape), dtype=np.float64)
    for fname in files.unique():
        sel = files == fname
        _, data = io.read_mrc(pathlib.Path(project, fname), mmap=True)
        mshape = data.shape if data.ndim == 2 else data.shape[1:]
        if tuple(shape) != tuple(mshape):
            raise ValueError(f"The shape of the MRC with filename {fname} differs from 'rlnImageSize'.")
        stack[np.nonzero(sel.to_numpy())[0]] = data if data.ndim == 2 else data[mrc_index[sel].to_numpy()]
    return stack


class RelionParticleStackDataset:
    """relion.py:536-862."""

    def __init__(self, parameter_file, path_to_relion_project, mode="r", mrcfile_settings=None, *,
                 loads_parameters=True):
        self._mode = _validate_mode(mode)
        self._parameter_file = parameter_file
        self._path_to_relion_project = pathlib.Path(path_to_relion_project)
        self.loads_parameters = loads_parameters
        self._mrcfile_settings = _mrcfile_settings(mrcfile_settings)
        particles = parameter_file.starfile_data["particles"]
        optics = parameter_file.starfile_data["optics"]
        if mode == "w":
            particles["rlnImageName"] = pd.Series(dtype=object)
            parameter_file.starfile_data = dict(optics=optics, particles=particles)
            self._path_to_relion_project.mkdir(parents=True, exist_ok=True)
        else:
            if "rlnImageName" not in particles.columns:
                raise OSError("Could not find column 'rlnImageName' in the STAR file; to write images set mode='w'.")
            if not self._path_to_relion_project.exists():
                raise FileNotFoundError("The RELION project directory `path_to_relion_project` does not exist.")

    @property
    def parameter_file(self):
        return self._parameter_file

    @property
    def mode(self):
        return self._mode

    @property
    def path_to_relion_project(self):
        return self._path_to_relion_project

    @property
    def mrcfile_settings(self):
        return self._mrcfile_settings

    @mrcfile_settings.setter
    def mrcfile_settings(self, value):
        self._mrcfile_settings = _mrcfile_settings(value)

    def __len__(self):
        return len(self._parameter_file)

    def __getitem__(self, index):
        pf = self._parameter_file
        particles = pf.starfile_data["particles"]
        _validate_dataset_index(type(self), index, len(particles))
        at = particles.iloc[index]
        frame = at.to_frame().T if isinstance(at, pd.Series) else at
        if "rlnImageName" not in frame.columns:
            raise OSError(f"No entry for 'rlnImageName' at index {index}.")
        optics = _optics_group_of(frame, pf.starfile_data["optics"])
        size = int(optics["rlnImageSize"])
        images = _load_image_stack_from_mrc((size, size), frame, self._path_to_relion_project)
        if isinstance(index, (int, np.integer)):
            images = images[0]
        parameters = pf[index] if self.loads_parameters else None
        return dict(parameters=parameters, images=images)

    def write_images(self, index_array, images, parameters=None):
        """relion.py:756-832: one MRCS file per call, rlnImageName filled in."""
        pf = self._parameter_file
        particles, optics = pf.starfile_data["particles"], pf.starfile_data["optics"]
        index_array = np.atleast_1d(np.asarray(index_array, dtype=np.int64))
        if parameters is None:
            group = _optics_group_of(particles.iloc[index_array], optics)
            pixel_size, dim = float(group["rlnImagePixelSize"]), int(group["rlnImageSize"])
        else:
            pixel_size = float(np.atleast_1d(parameters["image_config"].pixel_size)[0])
            dim = parameters["image_config"].shape[0]
        if not (images.ndim in (2, 3) and tuple(images.shape[-2:]) == (dim, dim)):
            raise ValueError(f"Image(s) must be of shape (n_images, {dim}, {dim}) or ({dim}, {dim}); got "
                             f"{tuple(images.shape)}.")
        if images.ndim == 2:
            images = images[None]
        if images.shape[0] != index_array.size:
            raise ValueError(f"Found {images.shape[0]} images for {index_array.size} dataset indices.")
        path, names = _make_image_filename(index_array, particles, len(pf), self._mrcfile_settings,
                                           self._path_to_relion_project)
        particles.loc[particles.index[index_array], "rlnImageName"] = names
        pf.starfile_data = dict(particles=particles, optics=optics)
        path.parent.mkdir(parents=True, exist_ok=True)
        try:
            io.write_image_stack_to_mrc(images, pixel_size, path, overwrite=self._mrcfile_settings["overwrite"],
                                        compression=self._mrcfile_settings["compression"])
        except Exception as err:
            raise OSError(f"Error writing the image stack to {path}: {err}") from err

    def __setitem__(self, index, value):
        if not isinstance(value, dict) or "images" not in value:
            raise TypeError("`dataset[index] = foo` needs a dict with key 'images' (and optionally 'parameters').")
        parameters = value.get("parameters")
        if parameters is not None:
            self._parameter_file[index] = parameters
        n = len(self._parameter_file)
        idx = np.asarray(range(*index.indices(n))) if isinstance(index, slice) else np.atleast_1d(index)
        self.write_images(idx, value["images"], parameters=parameters)

    def append(self, value):
        """relion.py:726-753."""
        if not isinstance(value, dict) or "images" not in value or value.get("parameters") is None:
            raise ValueError("`dataset.append(foo)` needs a dict with keys 'images' and 'parameters'.")
        start = len(self._parameter_file.starfile_data["particles"])
        self._parameter_file.append(value["parameters"])
        stop = len(self._parameter_file.starfile_data["particles"])
        self.write_images(np.arange(start, stop), value["images"], parameters=value["parameters"])


def _index_tree(index, tree):
    """Index the leading axis of every array leaf (dicts, lists, tuples, objects with __dict__)."""
    if tree is None or isinstance(tree, (str, bytes, int, float, bool)):
        return tree
    if isinstance(tree, dict):
        return {k: _index_tree(index, v) for k, v in tree.items()}
    if isinstance(tree, (list, tuple)):
        return type(tree)(_index_tree(index, v) for v in tree)
    if hasattr(tree, "shape") and hasattr(tree, "__getitem__") and getattr(tree, "ndim", 0) > 0:
        return tree[index]
    return tree


def simulate_particle_stack(dataset, compute_image_fn: Callable, constant_args: Any = None,
                            per_particle_args: Any = None, batch_size: int | None = None,
                            images_per_file: int | None = None, **kwargs):
    """particle_simulation.py:22-212.

    compute_image_fn(parameters, constant_args, per_particle_args) returns one
    image for one particle's parameters, or, with `batch_size`, a stack of
    images for a batch of particles: the parameter file returns batched leaves
    and the simulator forms a whole batch in one fused launch, which stands in
    for the reference's vmap over particles. Images are written one MRCS file
    per `images_per_file` particles, then the STAR file is saved
    (`kwargs` go to its save()).
    """
    if dataset.mode == "r":
        raise ValueError("The `dataset` is in reading mode (`mode = 'r'`); it must be opened with `mode = 'w'`.")
    n = len(dataset)
    per_file = n if images_per_file is None else int(images_per_file)
    pf = dataset.parameter_file
    for start in range(0, n, max(per_file, 1)):
        idx = np.arange(start, min(start + per_file, n), dtype=np.int64)
        images = _simulate_file(idx, pf, compute_image_fn, constant_args, per_particle_args, batch_size)
        dataset.write_images(idx, images, pf[idx])
    pf.save(**kwargs)


def _to_host(x):
    if hasattr(x, "detach"):
        return x.detach().to("cpu").numpy()
    return np.asarray(x)


def _simulate_file(idx, pf, fn, constant_args, per_particle_args, batch_size):
    if batch_size is None:
        imgs = []
        for i in idx:
            imgs.append(_to_host(fn(pf[int(i)], constant_args, _index_tree(int(i), per_particle_args))))
        return np.stack(imgs)
    out = []
    for b0 in range(0, idx.size, int(batch_size)):
        sub = idx[b0:b0 + int(batch_size)]
        res = fn(pf[sub], constant_args, _index_tree(sub, per_particle_args))
        out.append(_to_host(res).reshape(sub.size, *np.shape(res)[-2:]))
    return np.concatenate(out)
