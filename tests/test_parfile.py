from sem2dpack_amd import parfile


def test_namelist_parsing():
    p = parfile.parse_par("""# comment
&GENERAL iexec=1, ngll=5, fmax=1.d0 , ndof=2 ,
  title = 'a / b', verbose='1101' /
&MESH_CART xlim=-30d3,30d3, zlim=0.d0,1d4, nelem=120,60, ezflt=-1 /
&BC_DYNFLT friction='SWF', Tn=-120.d6,TtH='PWCONR', otd=0.2d0 /
&DIST_PWCONR num=2, ref=0d0,-7.5d3 / # Initial shear stress
     1.5d3
81.6d6 70.d6
&REC_LINE number = 11 , AtNode=F, first = -10d3,0d3 /
""")
    g = p.first("GENERAL")
    assert g.get("ngll") == 5 and g.get("fmax") == 1.0 and g.get("title") == "a / b"
    m = p.first("MESH_CART")
    assert m.get("xlim") == [-30e3, 30e3] and m.get("nelem") == [120, 60] and m.get("ezflt") == -1
    d = p.first("DIST_PWCONR")
    assert d.get("ref") == [0.0, -7.5e3] and d.data_lines == ["1.5d3", "81.6d6 70.d6"]
    assert p.first("REC_LINE").get("atnode") is False
    assert p.after(p.first("BC_DYNFLT"), "DIST_PWCONR") is d
