"""Copy INP meshes of the reference's test suite into tests/golden/inp/ as gzip
fixtures, so INP ingestion and the models built from them are tested without the
reference present: tests/element_test/C3D4.inp, C3D8.inp, C3D10.inp and
tests/pressure_test/C3D4Less.inp (the reference's Couple / reference-point model).

Run once where the reference is mounted:  python tests/golden/copy_inp_fixtures.py
"""
import gzip
import os
import shutil

SRC = '/root/reference/tests'
FILES = [('element_test', 'C3D4'), ('element_test', 'C3D8'), ('element_test', 'C3D10'),
         ('pressure_test', 'C3D4Less')]
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'inp')

if __name__ == '__main__':
    os.makedirs(DST, exist_ok=True)
    for sub, name in FILES:
        with open(os.path.join(SRC, sub, name + '.inp'), 'rb') as fi, \
                gzip.GzipFile(os.path.join(DST, name + '.inp.gz'), 'wb', mtime=0) as fo:
            shutil.copyfileobj(fi, fo)
        print('wrote', os.path.join(DST, name + '.inp.gz'))
