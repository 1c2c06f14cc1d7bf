"""Copy INP meshes of the reference's test suite into tests/golden/inp/ as gzip
fixtures, so INP ingestion and the models built from them are tested without the
reference present: tests/element_test/C3D4.inp, C3D8.inp, C3D10.inp,
tests/pressure_test/C3D4Less.inp (the reference's Couple / reference-point model) and
tests/gradient_test/C3D4.inp with its solution C3D4_results.npy (the reference's
self-contact + pressure + Couple model and its converged generalized coordinates, the
golden vector of tests/gradient_test/test_grad.py).

Run once where the reference is mounted:  python tests/golden/copy_inp_fixtures.py
"""
import gzip
import os
import shutil

import numpy as np

SRC = '/root/reference/tests'
FILES = [('element_test', 'C3D4'), ('element_test', 'C3D8'), ('element_test', 'C3D10'),
         ('pressure_test', 'C3D4Less'), ('gradient_test', 'C3D4')]
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'inp')

if __name__ == '__main__':
    os.makedirs(DST, exist_ok=True)
    for sub, name in FILES:
        out = name if sub != 'gradient_test' else 'gradient_' + name
        with open(os.path.join(SRC, sub, name + '.inp'), 'rb') as fi, \
                gzip.GzipFile(os.path.join(DST, out + '.inp.gz'), 'wb', mtime=0) as fo:
            shutil.copyfileobj(fi, fo)
        print('wrote', os.path.join(DST, out + '.inp.gz'))
    golden = np.load(os.path.join(SRC, 'gradient_test', 'C3D4_results.npy'), allow_pickle=False)
    np.save(os.path.join(os.path.dirname(DST), 'gradient_C3D4_results.npy'), golden)
    print('wrote gradient_C3D4_results.npy', golden.shape)
