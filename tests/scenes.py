"""Scene descriptions shared by the oracle and the library in the tests (the
generators live in the package: pbrtm/scenes.py)."""
from pbrtm.scenes import GREY, GREEN, RED, batch_jobs, box, cornell34, quad, soup_room, sphere_grid  # noqa: F401
