"""Command-line entry: `python -m voxeldvc_amd preprocess REF.npy DEF.npy
--output-dir DIR [...]` and `python -m voxeldvc_amd run WORK_DIR
[--l0-factor F]`, the `preprocess` and `run` subcommands of the reference's
CLI (src/voxeldvc/cli.py:13-61)."""
import sys


def main(argv=None):
    argv = list(sys.argv[1:] if argv is None else argv)
    if not argv or argv[0] in ('-h', '--help'):
        print(__doc__)
        return 0
    cmd, rest = argv[0], argv[1:]
    if cmd == 'run':
        from .run_dvc import main as run_main
        run_main(rest)
        return 0
    if cmd == 'preprocess':
        from .preprocess import main as pre_main
        pre_main(rest)
        return 0
    print('unknown subcommand %r (this package provides: preprocess, run)' % cmd,
          file=sys.stderr)
    return 2


if __name__ == '__main__':
    sys.exit(main())
