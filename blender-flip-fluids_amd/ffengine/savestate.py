"""Savestates of the marker particles in the add-on's on-disk format.

write_savestate() and load_savestate() follow the add-on's autosave path for
the state the FLIP substep keeps across frames, the marker particles
(src/addon/bake.py:2795-3100 writing, 554-722 and 771-777 loading):

- particle data go to raw packed float32 xyz files, marker_particle_position.data,
  marker_particle_velocity.data and, with the APIC transfer,
  marker_particle_affine{x,y,z}.data, written through the chunked range
  getters 2**21 particles at a time;
- autosave.state is the JSON record of the grid, the frame and the file names
  (sort_keys, indent 4), the attribute files this path does not produce named
  by empty strings as the add-on does when an attribute is off;
- loading reads the files back in chunks of 2**21 particles, one
  load_marker_particle_data (and load_marker_particle_affine_data) batch per
  chunk, then sets the current frame to frame_id + 1.

The functions take any object with the reference's FluidSimulation Python
surface, so a savestate written by one engine loads into the other.
"""
import json
import os

PARTICLES_PER_CHUNK = 2 ** 21
STATE_FILE = "autosave.state"
POSITION_FILE = "marker_particle_position.data"
VELOCITY_FILE = "marker_particle_velocity.data"
AFFINE_FILES = ("marker_particle_affinex.data", "marker_particle_affiney.data", "marker_particle_affinez.data")
_UNUSED_KEYS = ("marker_particle_age_filedata", "marker_particle_lifetime_filedata",
                "marker_particle_color_filedata", "marker_particle_source_id_filedata",
                "marker_particle_uid_filedata", "marker_particle_viscosity_filedata",
                "marker_particle_density_filedata", "marker_particle_id_filedata",
                "diffuse_particle_position_filedata", "diffuse_particle_velocity_filedata",
                "diffuse_particle_lifetime_filedata", "diffuse_particle_type_filedata",
                "diffuse_particle_id_filedata")


def _chunks(n):
    """[start, end) ranges of the add-on's chunked reads and writes
    (bake.py:2898-2903, 655-668): n // 2**21 + 1 chunks, the last one short."""
    for i in range(n // PARTICLES_PER_CHUNK + 1):
        yield i * PARTICLES_PER_CHUNK, min((i + 1) * PARTICLES_PER_CHUNK, n)


def write_savestate(fluidsim, directory, frameno, frame_start=0, frame_end=None):
    """Write the marker particles and autosave.state of `fluidsim` after frame
    `frameno` into `directory` (created if absent)."""
    os.makedirs(directory, exist_ok=True)
    isize, jsize, ksize = fluidsim.get_grid_dimensions()
    apic = fluidsim.is_velocity_transfer_method_APIC()
    n = fluidsim.get_num_marker_particles()
    getters = [("get_marker_particle_position_data_range", POSITION_FILE),
               ("get_marker_particle_velocity_data_range", VELOCITY_FILE)]
    if apic:
        getters += [("get_marker_particle_affine%s_data_range" % c, f) for c, f in zip("xyz", AFFINE_FILES)]
    for getter, name in getters:
        path = os.path.join(directory, name)
        with open(path, "wb") as f:
            for start, end in _chunks(n):
                f.write(getattr(fluidsim, getter)(start, end))
    info = {
        "isize": isize, "jsize": jsize, "ksize": ksize, "dx": fluidsim.get_cell_size(),
        "frame": frameno, "frame_start": frame_start,
        "frame_end": frameno if frame_end is None else frame_end,
        "frame_id": fluidsim.get_current_frame() - 1,
        "last_frame_id": (frameno if frame_end is None else frame_end) - frame_start,
        "num_marker_particles": n,
        "marker_particle_position_filedata": POSITION_FILE,
        "marker_particle_velocity_filedata": VELOCITY_FILE,
        "num_diffuse_particles": 0,
        "marker_particle_affinex_filedata": AFFINE_FILES[0] if apic else "",
        "marker_particle_affiney_filedata": AFFINE_FILES[1] if apic else "",
        "marker_particle_affinez_filedata": AFFINE_FILES[2] if apic else "",
    }
    for key in _UNUSED_KEYS:
        info[key] = ""
    with open(os.path.join(directory, STATE_FILE), "w", encoding="utf-8") as f:
        f.write(json.dumps(info, sort_keys=True, indent=4))
    return info


def read_savestate_info(directory):
    with open(os.path.join(directory, STATE_FILE), "r", encoding="utf-8") as f:
        return json.loads(f.read())


def load_savestate(fluidsim, directory):
    """Queue the savestate's marker particles into `fluidsim` (before its
    initialize) and set its current frame; returns the autosave record."""
    info = read_savestate_info(directory)
    n = int(info["num_marker_particles"])
    apic = fluidsim.is_velocity_transfer_method_APIC() and all(
        info.get("marker_particle_affine%s_filedata" % c) for c in "xyz")
    if n > 0:
        names = [info["marker_particle_position_filedata"], info["marker_particle_velocity_filedata"]]
        if apic:
            names += [info["marker_particle_affine%s_filedata" % c] for c in "xyz"]
        files = [open(os.path.join(directory, name), "rb") for name in names]
        try:
            for start, end in _chunks(n):
                count = end - start
                data = []
                for f in files:
                    f.seek(12 * start)
                    data.append(f.read(12 * count))
                fluidsim.load_marker_particle_data(count, data[0], data[1])
                if apic:
                    fluidsim.load_marker_particle_affine_data(count, data[2], data[3], data[4])
        finally:
            for f in files:
                f.close()
    fluidsim.set_current_frame(int(info["frame_id"]) + 1)
    return info
