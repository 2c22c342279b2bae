"""Material library (wakis/materials.py:27-44): key -> [eps_r, mu_r(, sigma)].

Keys are lower case; eps = eps_r*eps_0, mu = mu_r*mu_0; a third entry makes
the material conductive (sigma in S/m).
"""

import numpy as np

material_lib = {
    "pec": [np.inf, 1.0],
    "vacuum": [1.0, 1.0],
    "dielectric": [10.0, 1.0],
    "lossy metal": [10, 1.0, 10],
    "copper": [5.8e07, 1.0, 5.8e07],
    "berillium": [2.5e07, 1.0, 2.5e07],
    "stainless steel": [1.4e06, 1.0, 1.4e06],
}
