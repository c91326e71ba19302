"""Copies the reference's golden data files into tests/golden/ as fixtures.

Fixtures are data only: the NIST Standard Reference Simulation configurations
(public NIST data; sha256 as listed in the reference's data/nist/README.md,
checked here), the reference's committed exp01/exp02 result CSVs (the
108-atom NVE trajectories and the dt^2 summary), and the NIST published
energy tables as JSON. Run in a container that has the reference mounted:

    python tests/golden/make_fixtures.py [/root/reference]
"""
import hashlib
import json
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))

SHA256 = {
    "lj/lj_sample_config_periodic1.txt": "6fa4bf9d4cb84c07d4766e4c1143c5fdf0c560c47a695336c6d06b19e425d295",
    "lj/lj_sample_config_periodic2.txt": "d6714977c486280577f2fe69fdfd52188b47a1bdace653edf47ad44b29f37c9a",
    "lj/lj_sample_config_periodic3.txt": "c8e13fef51b21cea5f958ae911c0f45fe3f7ef4f92578d68569258bbb29113ec",
    "lj/lj_sample_config_periodic4.txt": "aa823314ac4abf597cbe9faf9b9caab267a78ddab1433779a6d7da72f59e1883",
    "spce/spce_sample_config_periodic1.txt": "a2b01a73b918444cd9a6a47895999924bb87b45b9315cf836710d7d5e2d840d6",
    "spce/spce_sample_config_periodic2.txt": "a52db516d5c8450f08cc71d220bfda26b4936596ba17210dd04297421da8238b",
    "spce/spce_sample_config_periodic3.txt": "5d89c3e07199c2fffff11076b35f5b40d3dce20e2dbad128719216634b0bcf92",
    "spce/spce_sample_config_periodic4.txt": "095a4ee6a64f281ebc0b77374e410e7dc2338e6bd0963643f74001b2946e68ee",
}

RESULTS = [
    "exp01_nist_table.csv",
    "exp02_energy_conservation.csv",
    "exp02_thermo_dt0.001.csv",
    "exp02_thermo_dt0.002.csv",
    "exp02_thermo_dt0.004.csv",
    "exp02_thermo_dt0.008.csv",
]

# NIST published values (reduced LJ units; SPC/E energies E/kB in K).
NIST = {
    "lj": [  # config, scheme, rc, Upair, Wpair, ULRC (5 significant figures)
        [1, "LRC", 3.0, -4.3515e3, -5.6867e2, -1.9849e2],
        [1, "LRC", 4.0, -4.4675e3, -1.2639e3, -8.3769e1],
        [1, "LFS", 3.0, -3.8709e3, 3.1754e2, 0.0],
        [2, "LRC", 3.0, -6.9000e2, -5.6846e2, -2.4230e1],
        [2, "LRC", 4.0, -7.0460e2, -6.5599e2, -1.0226e1],
        [2, "LFS", 3.0, -6.2012e2, -4.4533e2, 0.0],
        [3, "LRC", 3.0, -1.1467e3, -1.1649e3, -4.9622e1],
        [3, "LRC", 4.0, -1.1754e3, -1.3371e3, -2.0942e1],
        [3, "LFS", 3.0, -1.0210e3, -9.3578e2, 0.0],
        [4, "LRC", 3.0, -1.6790e1, -4.6249e1, -5.4517e-1],
        [4, "LRC", 4.0, -1.7060e1, -4.7869e1, -2.3008e-1],
        [4, "LFS", 3.0, -1.5001e1, -4.3096e1, 0.0],
    ],
    "spce": [  # config, Edisp, ELRC, Ereal, Efourier, Eself, Eintra, Etotal (6 s.f.)
        [1, 9.95387e4, -8.23715e2, -5.58889e5, 6.27009e3, -2.84469e6, 2.80999e6, -4.88604e5],
        [2, 1.93712e5, -3.29486e3, -1.19295e6, 6.03495e3, -5.68938e6, 5.61998e6, -1.06590e6],
        [3, 3.54344e5, -7.41343e3, -1.96297e6, 5.24461e3, -8.53407e6, 8.42998e6, -1.71488e6],
        [4, 4.48593e5, -1.37286e4, -3.57226e6, 7.58785e3, -1.42235e7, 1.41483e7, -3.20501e6],
    ],
    "coulomb_prefactor_K_A": 167100.956632,
    "lrc_hand": [  # rho, rc, U_LRC/N, P_LRC (4 significant figures)
        [0.8442, 2.5, -0.4520, -0.7621],
        [0.8442, 3.0, -0.2618, -0.4419],
        [0.7000, 2.5, -0.3748, -0.5240],
        [0.5000, 2.5, -0.2677, -0.2674],
    ],
}


def main(ref_root: str) -> None:
    for rel, digest in SHA256.items():
        src = os.path.join(ref_root, "data", "nist", rel)
        data = open(src, "rb").read()
        got = hashlib.sha256(data).hexdigest()
        if got != digest:
            raise SystemExit(f"sha256 mismatch for {rel}: {got}")
        dst = os.path.join(HERE, "nist", rel)
        os.makedirs(os.path.dirname(dst), exist_ok=True)
        open(dst, "wb").write(data)
    os.makedirs(os.path.join(HERE, "results"), exist_ok=True)
    for name in RESULTS:
        shutil.copyfile(os.path.join(ref_root, "results", name), os.path.join(HERE, "results", name))
    with open(os.path.join(HERE, "nist_values.json"), "w") as fh:
        json.dump(NIST, fh, indent=1)
    print("fixtures written under", HERE)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
