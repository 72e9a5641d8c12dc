"""The generated straight-line kinetics (phd_amd/csrc/mech_gen.h) is what
the generator writes for the bundled mechanism."""


def test_generated_file_is_current():
    from phd_amd import chemdata as cd
    from phd_amd.codegen_mech import OUT, generate
    with open(OUT) as f:
        assert f.read() == generate(cd.read_mechanism())
