from larmor_amd.cli import main

main()
